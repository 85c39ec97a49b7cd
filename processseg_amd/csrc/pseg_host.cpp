// Host side of libpseg: the plan compiler (float64 design and tables, as the
// reference builds its plans) and the C ABI of include/pseg.h.
//
// A compiled plan is one device allocation: the DevPlan (or DevCorr) the
// kernels read, followed by its tables (stage twiddles, split table, gains,
// pad tapers, AGC tables, the pilot multiplier), built on the host and copied
// once.  apply only launches kernels: it allocates nothing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pseg.h"
#include "pseg_plan.hpp"

using namespace pseg;

namespace {

constexpr double kPi = 3.141592653589793;
constexpr double kInf = std::numeric_limits<double>::infinity();

// ------------------------------------------------------------------ errors

thread_local std::string g_error;

pseg_status fail(pseg_status st, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_error = buf;
  return st;
}

pseg_status hip_fail(hipError_t e, const char *what) {
  return fail(PSEG_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

struct Invalid {
  std::string msg;
};

[[noreturn]] void invalid(const char *fmt, ...) {
  char buf[400];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  throw Invalid{buf};
}

bool finite_positive(double x) { return std::isfinite(x) && x > 0; }

template <typename T>
bool fits_in(double x) {
  return finite_positive(x) && finite_positive((double)(T)x);
}

// ---------------------------------------------------------------- lengths

bool is_fast(int64_t n) {
  if (n < 1) return false;
  for (int p : {2, 3, 5})
    while (n % p == 0) n /= p;
  return n == 1;
}

int64_t next_fast(int64_t n, bool even) {
  int64_t r = std::max<int64_t>(n, 1);
  while (!is_fast(r) || (even && (r & 1))) r++;
  return r;
}

// Stage radices of a complex length: 4s, then a 2, then 3s and 5s
// (fft/plan.nim:120-132).
std::vector<int> radices_of(int64_t m) {
  std::vector<int> r;
  while (m % 4 == 0) { r.push_back(4); m /= 4; }
  if (m % 2 == 0) { r.push_back(2); m /= 2; }
  while (m % 3 == 0) { r.push_back(3); m /= 3; }
  while (m % 5 == 0) { r.push_back(5); m /= 5; }
  return r;
}

FastDiv make_fastdiv(uint32_t d) {
  FastDiv f;
  f.d = d;
  uint32_t l = 0;
  while ((1ull << l) < d) l++;
  f.shift = l;
  f.mul = (uint32_t)((((1ull << 32) * ((1ull << l) - d)) / d) + 1);
  return f;
}

// ------------------------------------------------- float64 host transform
//
// A recursive decimation-in-time complex FFT over std::complex<double>, for
// plan building only (impulse lengths, the pilot spectrum).  Twiddles come
// from one table of exp(-2 pi i k / N) indexed exactly (k mod N).

using cd = std::complex<double>;

struct HostFft {
  int64_t N;
  std::vector<cd> W;
  explicit HostFft(int64_t n) : N(n), W(n) {
    for (int64_t k = 0; k < n; k++) {
      double a = -2 * kPi * (double)k / (double)n;
      W[k] = cd(std::cos(a), std::sin(a));
    }
  }
  void rec(cd *out, const cd *in, int64_t n, int64_t stride, int64_t wstep) const {
    if (n == 1) {
      out[0] = in[0];
      return;
    }
    int r = n % 4 == 0 ? 4 : n % 2 == 0 ? 2 : n % 3 == 0 ? 3 : 5;
    int64_t q = n / r;
    for (int j = 0; j < r; j++) rec(out + j * q, in + j * stride, q, stride * r, wstep * r);
    cd tmp[5];
    for (int64_t k = 0; k < q; k++) {
      for (int j = 0; j < r; j++) tmp[j] = out[j * q + k];
      for (int t = 0; t < r; t++) {
        cd acc = 0;
        for (int j = 0; j < r; j++) {
          int64_t e = ((int64_t)j * (k + t * q)) % n;
          acc += tmp[j] * W[(e * wstep) % N];
        }
        out[t * q + k] = acc;
      }
    }
  }
  // forward: X[k] = sum x[j] exp(-2 pi i jk / N)
  std::vector<cd> forward(const std::vector<cd> &x) const {
    std::vector<cd> y(N);
    rec(y.data(), x.data(), N, 1, 1);
    return y;
  }
};

// irfft of a real, even spectrum given by bins 0 .. n/2 (imaginary parts 0),
// scaled by 1/n.
std::vector<double> inverse_real(const std::vector<double> &re, int64_t n) {
  HostFft f(n);
  std::vector<cd> X(n);
  for (int64_t k = 0; k < n; k++) {
    int64_t kk = k <= n / 2 ? k : n - k;
    X[k] = cd(re[kk], 0.0);  // conj of a real value: itself
  }
  std::vector<cd> y = f.forward(X);  // conj(fft(conj(X))) of a real even X
  std::vector<double> h(n);
  for (int64_t j = 0; j < n; j++) h[j] = y[j].real() / (double)n;
  return h;
}

// ------------------------------------------------------------ filter design

bool has_low(const pseg_filter_spec &s) { return s.f2 > 0; }
bool has_high(const pseg_filter_spec &s) { return s.f3 < kInf; }

double warp(double f, double dt) {
  double x = kPi * f * dt / 1000;
  return x >= kPi / 2 ? kInf : std::tan(x);
}

double ramp(double x) {
  if (x <= 0) return 0.0;
  if (x >= 1) return 1.0;
  double s = std::sin(kPi / 2 * x);
  return s * s;
}

// response (filter/design.nim:236-255)
double response(const pseg_filter_spec &s, double f, double dt) {
  double r = 1.0;
  if (s.kind == 0) {
    if (has_low(s)) r *= ramp((f - s.f1) / (s.f2 - s.f1));
    if (has_high(s)) r *= 1 - ramp((f - s.f3) / (s.f4 - s.f3));
  } else {
    double wf = warp(f, dt);
    if (s.low_hz > 0)
      r *= f <= 0 ? 0.0 : 1 / (1 + std::pow(warp(s.low_hz, dt) / wf, s.low_slope / 6));
    if (s.high_hz < kInf)
      r *= wf == kInf ? 0.0 : 1 / (1 + std::pow(wf / warp(s.high_hz, dt), s.high_slope / 6));
  }
  for (int i = 0; i < s.n_notch; i++)
    r *= ramp(std::fabs(f - s.notch_hz[i]) / (s.notch_width[i] / 2));
  return r;
}

// narrowestRampHz (filter/design.nim:267-282)
double narrowest(const pseg_filter_spec &s) {
  double r = kInf;
  if (s.kind == 0) {
    if (has_low(s)) r = std::min(r, s.f2 - s.f1);
    if (has_high(s)) r = std::min(r, s.f4 - s.f3);
  } else {
    if (s.low_hz > 0) r = std::min(r, s.low_hz * std::min(1.0, 24 / s.low_slope) / 2);
    if (s.high_hz < kInf) r = std::min(r, s.high_hz * std::min(1.0, 24 / s.high_slope) / 2);
  }
  for (int i = 0; i < s.n_notch; i++) r = std::min(r, s.notch_width[i] / 2);
  return r;
}

// validate (filter/design.nim:100-173)
void validate(const pseg_filter_spec &s, double dt) {
  if (!(dt > 0 && dt < kInf)) invalid("dtMs must be finite and > 0, got %g", dt);
  if (s.kind != 0 && s.kind != 1) invalid("unknown response kind %d", s.kind);
  if (s.n_notch < 0 || s.n_notch > PSEG_MAX_NOTCH) invalid("at most %d notches", PSEG_MAX_NOTCH);
  double nyq = 500 / dt;
  if (s.kind == 0) {
    for (double f : {s.f1, s.f2, s.f3, s.f4})
      if (std::isnan(f)) invalid("a corner is NaN");
    if (has_low(s) && !(s.f1 >= 0 && s.f1 < s.f2))
      invalid("a low cut needs 0 <= f1 < f2 (a sloped ramp), got f1 = %g, f2 = %g", s.f1, s.f2);
    if (!has_low(s) && (s.f1 != 0 || s.f2 != 0)) invalid("no low cut needs f1 = f2 = 0");
    if (has_high(s) && !(s.f3 < s.f4 && s.f4 < kInf))
      invalid("a high cut needs f3 < f4 < Inf (a sloped ramp), got f3 = %g, f4 = %g", s.f3, s.f4);
    if (!has_high(s) && s.f4 != kInf) invalid("no high cut needs f3 = f4 = Inf");
    if (has_low(s) && has_high(s) && !(s.f2 <= s.f3)) invalid("a band-pass needs f2 <= f3");
    if (has_high(s) && s.f4 > nyq)
      invalid("f4 = %g Hz is above Nyquist (%g Hz at %g ms)", s.f4, nyq, dt);
    if (has_low(s) && s.f2 >= nyq) invalid("f2 = %g Hz is not below Nyquist (%g Hz)", s.f2, nyq);
  } else {
    if (!(s.low_hz >= 0 && s.low_hz < kInf))
      invalid("a low-cut corner must be finite and >= 0 (0: none), got %g", s.low_hz);
    if (!(s.high_hz > 0)) invalid("a high-cut corner must be > 0 (Inf: none), got %g", s.high_hz);
    if (s.low_hz > 0 && !(s.low_slope > 0 && s.low_slope < kInf))
      invalid("a slope must be finite and > 0 dB/octave, got %g", s.low_slope);
    if (s.high_hz < kInf && !(s.high_slope > 0 && s.high_slope < kInf))
      invalid("a slope must be finite and > 0 dB/octave, got %g", s.high_slope);
    if (!(s.low_hz < s.high_hz))
      invalid("the low-cut corner must be below the high-cut one, got %g and %g", s.low_hz,
              s.high_hz);
    if (s.high_hz < kInf && s.high_hz >= nyq)
      invalid("the high-cut corner %g Hz is not below Nyquist (%g Hz)", s.high_hz, nyq);
    if (s.low_hz >= nyq) invalid("the low-cut corner %g Hz is not below Nyquist (%g Hz)", s.low_hz, nyq);
  }
  for (int i = 0; i < s.n_notch; i++) {
    if (!(s.notch_hz[i] > 0 && s.notch_hz[i] < kInf))
      invalid("a notch frequency must be finite and > 0, got %g", s.notch_hz[i]);
    if (!(s.notch_width[i] > 0 && s.notch_width[i] < kInf))
      invalid("a notch width must be finite and > 0, got %g", s.notch_width[i]);
    if (s.notch_hz[i] >= nyq)
      invalid("a notch at %g Hz is not below Nyquist (%g Hz)", s.notch_hz[i], nyq);
  }
}

// energyLength (filter/design.nim:288-298)
int64_t energy_length(const std::vector<double> &h, double centre) {
  const int64_t n = (int64_t)h.size();
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

constexpr int64_t kMaxGrid = 1 << 24;

// featureLength (filter/design.nim:300-325)
int64_t feature_length(const pseg_filter_spec &s, double dt) {
  double feature = narrowest(s);
  if (feature == kInf) return 0;
  double fs = 1000 / dt;
  double start = std::ceil(16 * fs / feature);
  int64_t n = start > (double)kMaxGrid ? kMaxGrid + 1
                                       : next_fast(std::max<int64_t>(64, (int64_t)start), true);
  for (;;) {
    if (n > kMaxGrid)
      invalid("the filter's impulse response is longer than %lld samples at %g ms "
              "(narrowest ramp %g Hz)",
              (long long)(kMaxGrid / 8), dt, feature);
    std::vector<double> re(n / 2 + 1);
    double df = 1000 / ((double)n * dt);
    for (int64_t k = 0; k <= n / 2; k++) re[k] = response(s, (double)k * df, dt);
    if (s.zero_dc) re[0] = 0;
    std::vector<double> h = inverse_real(re, n);
    int64_t L = std::max(energy_length(h, h[0]), energy_length(h, 1 - h[0]));
    if (L <= n / 8) return L;
    n = next_fast(2 * n, true);
  }
}

// impulseLength + features (filter/design.nim:327-366)
int64_t impulse_length_uncached(const pseg_filter_spec &s, double dt) {
  validate(s, dt);
  std::vector<pseg_filter_spec> feats;
  pseg_filter_spec flat;
  memset(&flat, 0, sizeof flat);
  flat.kind = 0;
  flat.f3 = flat.f4 = kInf;
  if (s.kind == 0) {
    if (has_low(s)) { pseg_filter_spec f = flat; f.f1 = s.f1; f.f2 = s.f2; feats.push_back(f); }
    if (has_high(s)) { pseg_filter_spec f = flat; f.f3 = s.f3; f.f4 = s.f4; feats.push_back(f); }
  } else {
    pseg_filter_spec b;
    memset(&b, 0, sizeof b);
    b.kind = 1;
    if (s.low_hz > 0) {
      pseg_filter_spec f = b;
      f.low_hz = s.low_hz; f.low_slope = s.low_slope; f.high_hz = kInf; f.high_slope = s.high_slope;
      feats.push_back(f);
    }
    if (s.high_hz < kInf) {
      pseg_filter_spec f = b;
      f.low_hz = 0; f.low_slope = s.low_slope; f.high_hz = s.high_hz; f.high_slope = s.high_slope;
      feats.push_back(f);
    }
  }
  for (int i = 0; i < s.n_notch; i++) {
    pseg_filter_spec f = flat;
    f.n_notch = 1;
    f.notch_hz[0] = s.notch_hz[i];
    f.notch_width[0] = s.notch_width[i];
    feats.push_back(f);
  }
  int64_t L = 0;
  for (auto &f : feats) L = std::max(L, feature_length(f, dt));
  return L;
}

// impulse_length is a pure function of (spec, dt) and costs float64 FFTs of
// up to 2^24 points: plans compiled for many window widths (a viewer's zoom
// windows) share one evaluation per spec.
int64_t impulse_length(const pseg_filter_spec &s, double dt) {
  pseg_filter_spec k = s;
  k.reserved = 0;
  for (int i = std::max(0, std::min(k.n_notch, PSEG_MAX_NOTCH)); i < PSEG_MAX_NOTCH; i++)
    k.notch_hz[i] = k.notch_width[i] = 0;
  std::string key((const char *)&k, sizeof k);
  key.append((const char *)&dt, sizeof dt);
  static std::mutex mu;
  static std::map<std::string, int64_t> memo;
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = memo.find(key);
    if (it != memo.end()) return it->second;
  }
  const int64_t L = impulse_length_uncached(s, dt);  // throws for an invalid spec
  std::lock_guard<std::mutex> g(mu);
  memo.emplace(key, L);
  return L;
}

// --------------------------------------------------------- device blobs

// A host image of one device allocation: structs and tables appended with
// 256-byte alignment; pointers inside are offsets until patched.
struct Blob {
  std::vector<char> bytes;
  size_t add(const void *p, size_t n) {
    size_t off = (bytes.size() + 255) / 256 * 256;
    bytes.resize(off + n);
    if (n) memcpy(bytes.data() + off, p, n);
    return off;
  }
  template <typename V>
  size_t add_vec(const std::vector<V> &v) {
    return add(v.data(), v.size() * sizeof(V));
  }
};

// Pointer fields hold blob offsets + 1 until the device base is known.
template <typename P>
void set_off(P &field, size_t off) {
  field = (P)(uintptr_t)(off + 1);
}
template <typename P>
void patch(P &field, char *base) {
  uintptr_t v = (uintptr_t)field;
  if (v) field = (P)(base + (v - 1));
}

// Stage twiddles and split table of a real FFT of n points (fft/plan.nim:
// 136-160): float64, stored as T.
// The passes of an m-point complex transform as the device runs them
// (consecutive stages fused while fusable): their butterfly sizes.
std::vector<int> fft_passes(int64_t m, int fuse) {
  std::vector<int> rad = radices_of(m), out;
  for (size_t i = 0; i < rad.size();) {
    const bool pair = i + 1 < rad.size() && fusable(rad[i], rad[i + 1], fuse);
    out.push_back(pair ? rad[i] * rad[i + 1] : rad[i]);
    i += pair ? 2 : 1;
  }
  return out;
}

// The first pass's butterfly size when a second pass reads its padded
// output, else 0.
int32_t first_pass_pad(int64_t m, int fuse) {
  std::vector<int> passes = fft_passes(m, fuse);
  return passes.size() >= 2 ? passes[0] : 0;
}

// LDS bytes of the in-place buffer of an n-point real transform in T.
size_t fft_lds_bytes(int64_t n, size_t esz, int fuse) {
  const int64_t m = n % 2 ? n : n / 2;
  const int32_t pad = first_pass_pad(m, fuse);
  return (size_t)(m + (pad ? m / pad : 0)) * 2 * esz;
}

// Whether `threads` threads run every pass of an m-point transform in place
// (no thread holds more than pass_iters_max butterflies of a pass).
bool fft_fits(int64_t m, int threads, int fuse, int budget) {
  for (int r : fft_passes(m, fuse)) {
    const int64_t butterflies = m / r;
    if ((butterflies + threads - 1) / threads > pass_iters_max(r, budget)) return false;
  }
  return true;
}

// The coefficients of combine_gain / combine_cplx (the split, the per-bin
// multiplier and the inverse's unsplit as one linear map per bin pair) from
// the multipliers in float64, rounded once to T.  n even, m = n / 2; mul holds
// m + 1 bins (the device's 1/2 and 1/n already folded in).
template <typename T>
std::vector<T> lin_coef_gain(const std::vector<double> &g, int64_t n) {
  const int64_t m = n / 2;
  std::vector<T> c(4 * (m / 2 + 1), (T)0);
  c[0] = (T)g[0];
  c[1] = (T)g[m];
  for (int64_t k = 1; k <= m / 2; k++) {
    const double th = 2 * kPi * (double)k / (double)n;
    const double sn = -std::sin(th), cs = std::cos(th);
    const double gk = g[k], gkk = g[m - k];
    c[4 * k] = (T)(gk * (2 + 2 * sn) + gkk * (2 - 2 * sn));
    c[4 * k + 1] = (T)(gkk * (2 + 2 * sn) + gk * (2 - 2 * sn));
    c[4 * k + 2] = (T)(2 * cs * (gk - gkk));
  }
  return c;
}

template <typename T>
std::vector<T> lin_coef_cplx(const std::vector<cd> &mul, int64_t n) {
  const int64_t m = n / 2;
  std::vector<T> c(8 * (m / 2 + 1), (T)0);
  c[0] = (T)mul[0].real();
  c[1] = (T)mul[m].real();
  const cd I(0, 1);
  for (int64_t k = 1; k <= m / 2; k++) {
    const double th = 2 * kPi * (double)k / (double)n;
    const double sn = -std::sin(th), cs = std::cos(th);
    const cd a = mul[k], b = mul[m - k];
    const cd al = a * (2 + 2 * sn) + std::conj(b) * (2 - 2 * sn);
    const cd be = 2 * cs * I * (a - std::conj(b));
    const cd al2 = b * (2 + 2 * sn) + std::conj(a) * (2 - 2 * sn);
    const cd be2 = -2 * cs * I * (b - std::conj(a));
    const cd v[4] = {al, be, al2, be2};
    for (int i = 0; i < 4; i++) {
      c[8 * k + 2 * i] = (T)v[i].real();
      c[8 * k + 2 * i + 1] = (T)v[i].imag();
    }
  }
  return c;
}

template <typename T>
void build_fft(int64_t n, DevFft &f, Blob &blob, int fuse) {
  memset(&f, 0, sizeof f);
  f.n = (int32_t)n;
  // an odd length runs as the n-point complex transform of (x, 0)
  f.m = (int32_t)(n % 2 ? n : n / 2);
  std::vector<int> rad = radices_of(f.m);
  if ((int)rad.size() > kMaxRadices) invalid("an FFT of %lld points has too many stages", (long long)n);
  f.nrad = (int32_t)rad.size();
  std::vector<T> tw;
  int64_t cur = f.m, s = 1;
  for (int i = 0; i < f.nrad; i++) {
    int r = rad[i];
    int64_t sub = cur / r;
    f.rad[i] = r;
    f.sub[i] = (int32_t)sub;
    f.tw_off[i] = (int32_t)(tw.size() / 2);
    f.sdiv[i] = make_fastdiv((uint32_t)s);
    for (int64_t p = 0; p < sub; p++)
      for (int k = 1; k < r; k++) {
        double a = -2 * kPi * (double)((p * k) % cur) / (double)cur;
        tw.push_back((T)std::cos(a));
        tw.push_back((T)std::sin(a));
      }
    cur = sub;
    s *= r;
  }
  if (tw.empty()) tw.assign(2, (T)0);
  f.pad = first_pass_pad(f.m, fuse);
  f.stride = f.m + (f.pad ? f.m / f.pad : 0);
  std::vector<T> sp(2 * (f.m + 1));
  for (int64_t k = 0; k <= f.m; k++) {
    double a = -2 * kPi * (double)k / (double)n;
    sp[2 * k] = (T)std::cos(a);
    sp[2 * k + 1] = (T)std::sin(a);
  }
  set_off(f.tw, blob.add_vec(tw));
  set_off(f.split, blob.add_vec(sp));
}

void patch_fft(DevFft &f, char *base) {
  patch(f.tw, base);
  patch(f.split, base);
}

// ---------------------------------------------------------------- IIR --

struct Sos {
  double b0, b1, b2, a1, a2, g;
};

Sos sos_of(double b0, double b1, double b2, double a1, double a2) {
  return Sos{b0, b1, b2, a1, a2, (b0 + b1 + b2) / (1 + a1 + a2)};
}

// butterSections (filter/iir.nim:113-148): a digital Butterworth high-pass
// (high) or low-pass of `order` poles, -3 dB at fc Hz, by the bilinear
// transform prewarped to the corner; conjugate pole pairs as biquads, one
// first-order section for an odd order.
void butter_sections(double fc, double dt, int order, bool high, std::vector<Sos> &out) {
  const double fs = 1000 / dt;
  const double wc = 2 * fs * std::tan(kPi * fc / fs);
  const double k2 = 2 * fs;
  for (int k = 0; k < order / 2; k++) {
    const double th = kPi * (double)(2 * k + 1) / (double)(2 * order);
    const std::complex<double> sp(-wc * std::sin(th), wc * std::cos(th));
    const std::complex<double> z = (std::complex<double>(k2, 0) + sp) / (std::complex<double>(k2, 0) - sp);
    const double a1 = -2 * z.real(), a2 = z.real() * z.real() + z.imag() * z.imag();
    const double g = high ? (1 - a1 + a2) / 4 : (1 + a1 + a2) / 4;
    out.push_back(high ? sos_of(g, -2 * g, g, a1, a2) : sos_of(g, 2 * g, g, a1, a2));
  }
  if (order % 2 == 1) {
    const double a1 = -(k2 - wc) / (k2 + wc);
    const double g = high ? (1 - a1) / 2 : (1 + a1) / 2;
    out.push_back(high ? sos_of(g, -g, 0, a1, 0) : sos_of(g, g, 0, a1, 0));
  }
}

// causalLength (filter/iir.nim:150-182): samples after a spike holding
// 99.99 % of the energy of the causal impulse response, or of its
// complement, whichever is longer.
int64_t causal_length(const std::vector<Sos> &secs) {
  for (int64_t n = 1024;; n *= 2) {
    std::vector<double> h(n, 0.0);
    h[0] = 1;
    for (const Sos &c : secs) {
      double s1 = 0, s2 = 0;
      for (int64_t j = 0; j < n; j++) {
        const double x = h[j];
        const double y = c.b0 * x + s1;
        s1 = c.b1 * x - c.a1 * y + s2;
        s2 = c.b2 * x - c.a2 * y;
        h[j] = y;
      }
    }
    double rest = 0, tail = 0;
    for (int64_t j = 1; j < n; j++) {
      rest += h[j] * h[j];
      if (j >= n - n / 4) tail += h[j] * h[j];
    }
    if (tail <= 1e-12 * std::max(rest, 1e-300) || n >= ((int64_t)1 << 24)) {
      int64_t best = 0;
      for (double c0 : {h[0], 1 - h[0]}) {
        const double total = c0 * c0 + rest;
        double e = c0 * c0;
        int64_t L = 0;
        while (e < 0.9999 * total && L < n - 1) {
          L++;
          e += h[L] * h[L];
        }
        best = std::max(best, L);
      }
      return best;
    }
  }
}

int iir_poles(double slope, double per_pole, const char *side) {
  const double n = slope / per_pole;
  if (!(n >= 1 && n <= 32 && n == std::round(n)))
    invalid("the %s slope of an IIR filter must be a multiple of %g dB/octave (%d per pole), "
            "1 .. 32 poles; got %g", side, per_pole, (int)per_pole, slope);
  return (int)n;
}

// initIirPlan (filter/iir.nim:193-239) up to the plan's fields: the
// sections, the margin (the spec's impulse length for zero phase, the
// causal length for minimum phase), G(1) and G'(1).
struct IirDesign {
  std::vector<Sos> secs;
  int64_t pad;
  double dc_gain, dc_slope;
};

IirDesign design_iir(const pseg_filter_spec &spec, double dt, int phase) {
  if (phase != 0 && phase != 1) invalid("unknown IIR phase %d", phase);
  validate(spec, dt);
  if (spec.kind != 1)
    invalid("an IIR filter needs a Butterworth spec (butterLowCut, butterHighCut, butterBandPass)");
  if (spec.n_notch > 0 || spec.zero_dc) invalid("an IIR filter takes no notches and no 0 Hz suppression");
  const double per_pole = phase == 0 ? 12.0 : 6.0;
  std::vector<Sos> low, high;
  if (spec.low_hz > 0) butter_sections(spec.low_hz, dt, iir_poles(spec.low_slope, per_pole, "low-cut"), true, low);
  if (spec.high_hz < kInf)
    butter_sections(spec.high_hz, dt, iir_poles(spec.high_slope, per_pole, "high-cut"), false, high);
  IirDesign d;
  d.secs = low;
  d.secs.insert(d.secs.end(), high.begin(), high.end());
  d.pad = phase == 0 ? impulse_length(spec, dt) : std::max(causal_length(low), causal_length(high));
  double g = 1, dg = 0;
  for (const Sos &c : d.secs) {
    const double n1 = c.b0 + c.b1 + c.b2, d1 = 1 + c.a1 + c.a2;
    const double gi = n1 / d1;
    const double dgi = ((c.b1 + 2 * c.b2) * d1 - n1 * (c.a1 + 2 * c.a2)) / (d1 * d1);
    dg = dg * gi + g * dgi;
    g *= gi;
  }
  d.dc_gain = phase == 0 ? g * g : g;
  d.dc_slope = phase == 0 ? 0.0 : dg;
  return d;
}

// gainDbAt / factorAt (gain/curves.nim:134-160) of a gainCurve op, for its
// table: linear in dB between knots (bisection), flat outside them,
// 10^(dB/20); (max(t, minMs)/refMs)^power directly for a power curve.
double gain_factor_host(const pseg_op &op, double t) {
  if (std::isnan(t)) return t;
  if (op.gain_kind == 1) return std::pow(std::fmax(t, op.gain_min_ms) / op.gain_ref_ms, op.gain_power);
  double db;
  if (op.gain_kind == 2) {
    db = op.gain_db_per_s * t / 1000;
  } else {
    const double *kt = op.gain_knot_ms, *kd = op.gain_knot_db;
    const int last = op.gain_knots - 1;
    if (t <= kt[0]) {
      db = kd[0];
    } else if (t >= kt[last]) {
      db = kd[last];
    } else {
      int lo = 0, hi = last;
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (kt[mid] <= t) lo = mid;
        else hi = mid;
      }
      const double f = (t - kt[lo]) / (kt[hi] - kt[lo]);
      db = kd[lo] + f * (kd[hi] - kd[lo]);
    }
  }
  return std::pow(10.0, db / 20);
}

const char *op_name(int kind) {
  static const char *names[] = {"opDemean", "opDetrend", "opScale", "opGainDb", "opClip",
                                "opPolarity", "opMuteTop", "opMuteBottom", "opFilter",
                                "opIir", "opAgc", "opNormalize", "opGainCurve",
                                "opCalibrate", "opUncalibrate"};
  return kind >= 0 && kind <= 14 ? names[kind] : "op?";
}

std::atomic<uint64_t> g_generation{0};

}  // namespace

// --------------------------------------------------------------- the plans

struct pseg_corr_plan {
  int dtype;
  double dt_ms, scale;
  int32_t nfft, shift, out, pilot, record;
  size_t lds;
  int threads;
  DevCorr *dev;
};

namespace {

template <typename T>
void compile_step(const pseg_op &op, double dt, int32_t samples, DevStep &st, Blob &blob,
                  pseg_plan &P, int i) {
  memset(&st, 0, sizeof st);
  P.step_kind[i] = op.kind;
  size_t lds = 0;
  switch (op.kind) {
    case PSEG_OP_DEMEAN: st.kind = kDemean; break;
    case PSEG_OP_DETREND: st.kind = kDetrend; break;
    case PSEG_OP_POLARITY: st.kind = kPolarity; break;
    case PSEG_OP_SCALE:
      if (!std::isfinite(op.value)) invalid("the factor must be finite, got %g", op.value);
      if (!std::isfinite((double)(T)op.value))
        invalid("a factor of %g overflows %s", op.value, sizeof(T) == 4 ? "float32" : "float64");
      st.kind = kScale;
      st.factor = op.value;
      break;
    case PSEG_OP_GAIN_DB: {
      if (!std::isfinite(op.value)) invalid("the gain must be finite, got %g dB", op.value);
      double f = std::pow(10.0, op.value / 20);
      if (!std::isfinite(f) || !std::isfinite((double)(T)f))
        invalid("a gain of %g dB overflows %s", op.value, sizeof(T) == 4 ? "float32" : "float64");
      st.kind = kScale;
      st.factor = f;
      break;
    }
    case PSEG_OP_CLIP:
      if (std::isnan(op.lo) || std::isnan(op.hi) || op.lo > op.hi)
        invalid("clip limits must satisfy lo <= hi, got %g .. %g", op.lo, op.hi);
      st.kind = kClip;
      st.lo = op.lo;
      st.hi = op.hi;
      break;
    case PSEG_OP_FILTER: {
      // initFilterPlan (filter/apply.nim:118-151)
      const pseg_filter_spec &spec = op.spec;
      if (op.trend < 0 || op.trend > 2) invalid("unknown trend removal %d", op.trend);
      validate(spec, dt);
      int64_t margin = impulse_length(spec, dt);
      double dc = response(spec, 0, dt);
      bool reflect = dc >= 0.5 || op.trend != 2;
      int64_t pad = reflect ? margin : std::min<int64_t>(margin, samples);
      int64_t n = samples + (reflect ? 4 : 2) * pad;
      double rmp = narrowest(spec);
      if (rmp < kInf && pad == margin) n = std::max<int64_t>(n, (int64_t)std::ceil(2000 / (rmp * dt)));
      n = next_fast(n, true);
      if (n > (1 << 20)) invalid("an FFT of %lld points is beyond this GPU path", (long long)n);
      DevFilter &F = st.filter;
      build_fft<T>(n, F.fft, blob, kFusePipeline);
      std::vector<double> gain(n / 2 + 1);
      double df = 1000 / ((double)n * dt);
      // the device folds the split's 1/2 and the inverse's 1/n into the gain
      const double fold = 0.5 / (double)n;
      for (int64_t k = 0; k <= n / 2; k++)
        gain[k] = response(spec, (double)k * df, dt) * fold;
      if (spec.zero_dc) gain[0] = 0;
      std::vector<T> fall(std::max<int64_t>(pad, 1));
      for (int64_t k = 1; k <= pad; k++) {
        double c = std::cos(kPi / 2 * (double)k / (double)(pad + 1));
        fall[k - 1] = (T)(c * c);
      }
      set_off(F.coef, blob.add_vec(lin_coef_gain<T>(gain, n)));
      set_off(F.fall, blob.add_vec(fall));
      F.pad = (int32_t)pad;
      F.reflect = reflect;
      F.zero_dc = spec.zero_dc != 0;
      F.trend = op.trend;
      F.dc_gain = dc;
      st.kind = kFilter;
      P.step_nfft[i] = (int32_t)n;
      P.step_pad[i] = (int32_t)pad;
      P.step_margin[i] = (int32_t)margin;
      P.step_reflect[i] = reflect;
      P.margin += (int32_t)margin;
      lds = fft_lds_bytes(n, sizeof(T), kFusePipeline);
      break;
    }
    case PSEG_OP_AGC: {
      // initAgcPlan (gain/gain.nim:272-302)
      if (op.mode != 0 && op.mode != 1) invalid("unknown AGC mode %d", op.mode);
      if (!finite_positive(op.window_ms))
        invalid("an AGC window must be finite and positive, got %g ms", op.window_ms);
      double h = std::round(op.window_ms / (2 * dt));
      if (h < 1) invalid("an AGC window of %g ms is under 3 samples at %g ms", op.window_ms, dt);
      if (h > (double)(1 << 26)) invalid("an AGC window of %g ms is too long", op.window_ms);
      if (!finite_positive(op.target)) invalid("an AGC target must be finite and positive, got %g", op.target);
      if (!finite_positive(op.epsilon)) invalid("an AGC epsilon must be finite and positive, got %g", op.epsilon);
      int64_t H = (int64_t)h;
      std::vector<double> inv(2 * H + 2), rev(2 * H + 1);
      inv[0] = 0;
      for (int64_t c = 1; c < 2 * H + 2; c++) inv[c] = 1 / (double)c;
      for (int64_t m = 0; m < 2 * H + 1; m++) rev[m] = 1 / (double)(2 * H + 1 - m);
      DevAgc &G = st.agc;
      G.half = (int32_t)H;
      G.mode = op.mode;
      G.target = op.target;
      G.epsilon = op.epsilon;
      set_off(G.inv, blob.add_vec(inv));
      set_off(G.rev, blob.add_vec(rev));
      st.kind = kAgc;
      P.step_margin[i] = (int32_t)H;
      P.margin += (int32_t)H;
      // x / suffix sums / interior means, plus the head's prefix sums and P[n]
      lds = ((size_t)samples + (size_t)std::min<int64_t>(H, samples) + 1) * sizeof(T);
      break;
    }
    case PSEG_OP_NORMALIZE: {
      // initNormalizePlan (gain/gain.nim:153-167)
      if (op.stat < 0 || op.stat > 2) invalid("unknown statistic %d", op.stat);
      if (!finite_positive(op.target))
        invalid("a normalize target must be finite and positive, got %g", op.target);
      if (!(op.percentile >= 0 && op.percentile <= 100))
        invalid("a percentile must be in 0 .. 100, got %g", op.percentile);
      st.kind = op.scope == 1 ? kNormPanel : kNormTrace;
      st.stat = op.stat;
      st.target = op.target;
      st.percentile = op.percentile;
      P.norm_stat[i] = op.stat;
      P.norm_target[i] = op.target;
      P.norm_pct[i] = op.percentile;
      lds = 256 * sizeof(uint32_t);
      break;
    }
    case PSEG_OP_MUTE_TOP:
    case PSEG_OP_MUTE_BOTTOM: {
      // muteTopOp / muteBottomOp (pipeline.nim:417-424) with the curve
      // checks of timeMute / tableMute / lmoMute (core/mute.nim:58-98)
      if (!std::isfinite(op.taper_ms) || op.taper_ms < 0)
        invalid("a mute taper must be finite and >= 0 ms, got %g", op.taper_ms);
      DevMute &M = st.mute;
      M.kind = op.mute_kind;
      M.taper_ms = op.taper_ms;
      M.dt_ms = dt;
      if (op.mute_kind == 0) {
        if (!std::isfinite(op.mute_t_ms)) invalid("a mute time must be finite, got %g", op.mute_t_ms);
        M.t_ms = op.mute_t_ms;
      } else if (op.mute_kind == 2) {
        if (!finite_positive(op.mute_ms_per_dist))
          invalid("an LMO mute velocity must be finite and positive");
        if (!std::isfinite(op.mute_t_ms) || !std::isfinite(op.mute_pad_ms))
          invalid("LMO mute times must be finite, got t0Ms %g, padMs %g", op.mute_t_ms, op.mute_pad_ms);
        M.t_ms = op.mute_t_ms;
        M.ms_per_dist = op.mute_ms_per_dist;
        M.pad_ms = op.mute_pad_ms;
      } else if (op.mute_kind == 1) {
        if (op.mute_knots < 1 || op.mute_knots > PSEG_MAX_KNOTS)
          invalid("a mute table needs 1 .. %d knots, got %d", PSEG_MAX_KNOTS, op.mute_knots);
        std::vector<double> o(op.mute_offsets, op.mute_offsets + op.mute_knots);
        std::vector<double> t(op.mute_times, op.mute_times + op.mute_knots);
        for (int32_t k = 0; k < op.mute_knots; k++) {
          if (!std::isfinite(o[k]) || o[k] < 0) invalid("mute table offsets must be finite and >= 0, got %g", o[k]);
          if (k > 0 && o[k] <= o[k - 1]) invalid("mute table offsets must be strictly ascending: %g is followed by %g", o[k - 1], o[k]);
          if (!std::isfinite(t[k])) invalid("mute table times must be finite, got %g", t[k]);
        }
        M.knots = op.mute_knots;
        set_off(M.offsets, blob.add_vec(o));
        set_off(M.times, blob.add_vec(t));
      } else {
        invalid("unknown mute curve %d", op.mute_kind);
      }
      st.kind = op.kind == PSEG_OP_MUTE_TOP ? kMuteTop : kMuteBottom;
      lds = 0;
      break;
    }
    case PSEG_OP_GAIN_CURVE: {
      // gainCurveOp (pipeline.nim:440-442): initGainCurvePlan (gain/curves.
      // nim:214-249) tabulated to the end of the record, with the curve
      // checks of knotGain / powerGain / exponentialGain (gain/curves.nim:
      // 100-131)
      DevGain &G = st.gain;
      G.kind = op.gain_kind;
      if (op.gain_kind == 1) {
        if (!std::isfinite(op.gain_power)) invalid("a power must be finite, got %g", op.gain_power);
        if (!finite_positive(op.gain_ref_ms)) invalid("refMs must be finite and positive, got %g", op.gain_ref_ms);
        if (!finite_positive(op.gain_min_ms)) invalid("minMs must be finite and positive, got %g", op.gain_min_ms);
      } else if (op.gain_kind == 2) {
        if (!std::isfinite(op.gain_db_per_s)) invalid("dbPerS must be finite, got %g", op.gain_db_per_s);
      } else if (op.gain_kind == 0) {
        if (op.gain_knots < 1 || !op.gain_knot_ms || !op.gain_knot_db)
          invalid("the curve has no knots: build it with knotGain, powerGain or exponentialGain");
        if (op.gain_knots > PSEG_MAX_GAIN_KNOTS)
          invalid("a knot gain of %d knots (at most %d); UNSUPPORTED", op.gain_knots, PSEG_MAX_GAIN_KNOTS);
        for (int32_t k = 0; k < op.gain_knots; k++) {
          if (!std::isfinite(op.gain_knot_ms[k]) || !std::isfinite(op.gain_knot_db[k]))
            invalid("knot %d is not finite: (%g ms, %g dB)", k, op.gain_knot_ms[k], op.gain_knot_db[k]);
          if (k > 0 && !(op.gain_knot_ms[k] > op.gain_knot_ms[k - 1]))
            invalid("knot times must strictly increase: %g ms is followed by %g ms",
                    op.gain_knot_ms[k - 1], op.gain_knot_ms[k]);
        }
      } else {
        invalid("unknown gain curve %d", op.gain_kind);
      }
      const double lo = op.gain_tmin_ms;
      const double hi = std::isnan(op.gain_tmax_ms) ? lo + (double)(P.record - 1) * dt : op.gain_tmax_ms;
      if (!(std::isfinite(lo) && std::isfinite(hi)))
        invalid("the table limits must be finite, got %g .. %g", lo, hi);
      if (hi < lo) invalid("tMaxMs %g is before tMinMs %g", hi, lo);
      const double span = std::floor((hi - lo) / dt + 1e-9);
      if (span >= (double)(1 << 26)) invalid("a gain table of %g samples is too long", span);
      G.len = (int64_t)span + 1;
      G.power = op.gain_power;
      G.ref_ms = op.gain_ref_ms;
      G.min_ms = op.gain_min_ms;
      G.db_per_s = op.gain_db_per_s;
      G.t_min_ms = lo;
      G.dt_ms = dt;
      std::vector<T> tab((size_t)G.len);
      for (int64_t k = 0; k < G.len; k++) {
        const double t = lo + (double)k * dt;
        const double g = gain_factor_host(op, t);
        tab[k] = (T)g;
        // 0 · Inf would turn muted zeros into NaN
        if (std::isinf((double)tab[k]) || std::isnan(g))
          invalid("the gain at %g ms (%g) overflows %s", t, g, sizeof(T) == 4 ? "float32" : "float64");
      }
      set_off(G.table, blob.add_vec(tab));
      if (op.gain_kind == 0) {
        G.knots = op.gain_knots;
        set_off(G.knot_ms, blob.add_vec(std::vector<double>(op.gain_knot_ms, op.gain_knot_ms + op.gain_knots)));
        set_off(G.knot_db, blob.add_vec(std::vector<double>(op.gain_knot_db, op.gain_knot_db + op.gain_knots)));
      }
      st.kind = kGainCurve;
      lds = (size_t)samples * sizeof(T);  // the row's factors
      break;
    }
    case PSEG_OP_IIR: {
      // iirOp (pipeline.nim:428-430): initIirPlan (filter/iir.nim:193-239)
      if (op.trend < 0 || op.trend > 2) invalid("unknown trend removal %d", op.trend);
      const IirDesign d = design_iir(op.spec, dt, op.iir_phase);
      DevIir &I = st.iir;
      const bool reflect = std::fabs(d.dc_gain) >= 0.5 || op.trend != 2;
      const int64_t lead = reflect ? d.pad : 0;
      const int64_t tail = op.iir_phase == 0 ? d.pad : 0;
      const int64_t span = lead + samples + tail;
      if (span > (1 << 20)) invalid("an IIR row of %lld padded samples is beyond this GPU path; UNSUPPORTED", (long long)span);
      I.n_sections = (int32_t)d.secs.size();
      I.zero_phase = op.iir_phase == 0;
      I.reflect = reflect;
      I.trend = op.trend;
      I.lead = (int32_t)lead;
      I.tail = (int32_t)tail;
      I.span = (int32_t)span;
      I.dc_gain = d.dc_gain;
      I.dc_slope = d.dc_slope;
      std::vector<double> sec;
      for (const Sos &c : d.secs) sec.insert(sec.end(), {c.b0, c.b1, c.b2, c.a1, c.a2, c.g});
      set_off(I.sec, blob.add_vec(sec));
      st.kind = kIir;
      P.step_margin[i] = (int32_t)d.pad;
      P.step_pad[i] = (int32_t)lead;
      P.step_reflect[i] = reflect;
      P.margin += (int32_t)d.pad;
      // the padded row in float64; the scan's six float64 per thread are
      // added once the plan's threads per row are known
      lds = (size_t)span * sizeof(double);
      break;
    }
    case PSEG_OP_CALIBRATE:
    case PSEG_OP_UNCALIBRATE: {
      // initCalibrationPlan(op.calibration) (pipeline.nim:444-447): the
      // rules are checked and indexed here; each apply resolves the rows'
      // factors (pseg_plan_calib_factors) and the step multiplies by them
      if (P.n_calib >= 2) invalid("at most one calibration and one uncalibration per pipeline");
      auto cp = calib_new();
      const std::string e = calib_compile(op.calib_rules, op.calib_n_rules, op.calib_target,
                                          op.calib_file_factors, op.calib_scale_aux, *cp);
      if (!e.empty()) throw Invalid{e};
      st.kind = kCalibrate;
      st.stat = P.n_calib;
      P.calib[P.n_calib] = cp;
      P.calib_inverse[P.n_calib] = op.kind == PSEG_OP_UNCALIBRATE;
      P.n_calib++;
      break;
    }
    default:
      invalid("unknown op kind %d", op.kind);
  }
  P.step_lds[i] = lds;
}

void patch_step(DevStep &st, char *base) {
  if (st.kind == kFilter) {
    patch_fft(st.filter.fft, base);
    patch(st.filter.coef, base);
    patch(st.filter.fall, base);
  } else if (st.kind == kAgc) {
    patch(st.agc.inv, base);
    patch(st.agc.rev, base);
  } else if (st.kind == kMuteTop || st.kind == kMuteBottom) {
    patch(st.mute.offsets, base);
    patch(st.mute.times, base);
  } else if (st.kind == kIir) {
    patch(st.iir.sec, base);
  } else if (st.kind == kGainCurve) {
    patch(st.gain.table, base);
    patch(st.gain.knot_ms, base);
    patch(st.gain.knot_db, base);
  }
}

template <typename T>
pseg_status compile_plan(const pseg_op *ops, int32_t n_ops, double dt, int32_t samples,
                         int32_t record, pseg_plan **out) {
  if (!std::isfinite(dt) || dt <= 0) return fail(PSEG_E_INVALID, "dtMs must be finite and positive, got %g", dt);
  if (samples < 1) return fail(PSEG_E_INVALID, "a pipeline needs samples >= 1, got %d", samples);
  if (record != 0 && record < samples)
    return fail(PSEG_E_INVALID,
                "recordSamples (%d) must be 0 or >= samples (%d): it is the length of the "
                "traces a window of `samples` is cut from", record, samples);
  if (n_ops < 0 || n_ops > PSEG_MAX_OPS) return fail(PSEG_E_INVALID, "at most %d ops", PSEG_MAX_OPS);
  if (n_ops > 0 && !ops) return fail(PSEG_E_INVALID, "ops is NULL");
  if (pipeline_threads_for(samples) == 0)
    return fail(PSEG_E_UNSUPPORTED, "rows of %d samples are beyond this GPU path (at most %d)",
                samples, pipeline_max_cols(512));
  pseg_plan *P = new pseg_plan();  // value-initialized: every plain field 0
  P->dtype = sizeof(T) == 4 ? PSEG_F32 : PSEG_F64;
  P->dt_ms = dt;
  P->samples = samples;
  P->record = std::max(samples, record);
  Blob blob;
  blob.bytes.resize(sizeof(DevPlan));
  DevPlan &H = P->host;
  memset(&H, 0, sizeof H);
  // the calibration's place (pipeline.nim:495-511): before every op that
  // changes amplitudes, one of each kind; and the output's amplitude units
  int changed_at = -1, seen_cal[2] = {-1, -1};
  for (int i = 0; i < n_ops; i++) {
    const int kind = ops[i].kind;
    if (kind == PSEG_OP_CALIBRATE || kind == PSEG_OP_UNCALIBRATE) {
      if (changed_at >= 0) {
        delete P;
        return fail(PSEG_E_INVALID, "op %d (%s): a calibration must come before every op that "
                    "changes amplitudes; op %d (%s) does", i, op_name(kind), changed_at,
                    op_name(ops[changed_at].kind));
      }
      int &seen = seen_cal[kind - PSEG_OP_CALIBRATE];
      if (seen >= 0) {
        const int first = seen;
        delete P;
        return fail(PSEG_E_INVALID, "op %d (%s): one per pipeline: op %d is one too", i,
                    op_name(kind), first);
      }
      seen = i;
      P->amplitude_units = kind == PSEG_OP_CALIBRATE ? 1 : 0;
    } else if (kind == PSEG_OP_SCALE || kind == PSEG_OP_GAIN_DB || kind == PSEG_OP_CLIP ||
               kind == PSEG_OP_AGC || kind == PSEG_OP_NORMALIZE || kind == PSEG_OP_GAIN_CURVE) {
      if (changed_at < 0) changed_at = i;
      if (P->amplitude_units == 1) P->amplitude_units = 2;
    }
    try {
      compile_step<T>(ops[i], dt, samples, H.steps[i], blob, *P, i);
      if (P->step_lds[i] > kMaxRowLds)
        invalid("%zu bytes of LDS per row (an FFT of %d points) do not fit one workgroup "
                "(at most %zu); UNSUPPORTED", P->step_lds[i], P->step_nfft[i], kMaxRowLds);
    } catch (const Invalid &e) {
      delete P;
      if (e.msg == "UNSUPPORTED")
        return fail(PSEG_E_UNSUPPORTED, "op %d (%s): not part of the GPU path; keep the CPU plan "
                    "for this flow", i, op_name(ops[i].kind));
      const std::string tag = "; UNSUPPORTED";
      if (e.msg.size() > tag.size() && e.msg.compare(e.msg.size() - tag.size(), tag.size(), tag) == 0)
        return fail(PSEG_E_UNSUPPORTED, "op %d (%s): %s", i, op_name(ops[i].kind),
                    e.msg.substr(0, e.msg.size() - tag.size()).c_str());
      return fail(PSEG_E_INVALID, "op %d (%s): %s", i, op_name(ops[i].kind), e.msg.c_str());
    }
  }
  H.n_steps = n_ops;
  P->n_steps = n_ops;
  for (int i = 0; i < n_ops; i++) P->full = std::max(P->full, step_form(H.steps[i].kind));
  // the fewest threads per row that hold the row (24 samples a thread) and
  // run every transform of the plan in place
  for (int t = pipeline_threads_for(samples); t != 0 && P->threads == 0; t *= 2) {
    if (t > 512) break;
    bool fits = true;
    for (int i = 0; i < n_ops; i++)
      if (P->step_nfft[i] > 0 && !fft_fits(P->step_nfft[i] / 2, t, kFusePipeline, kBudgetPipeline)) fits = false;
    if (fits) P->threads = t;
  }
  if (P->threads == 0) {
    delete P;
    return fail(PSEG_E_UNSUPPORTED, "the plan's FFTs are too long for the in-place transform "
                "of one workgroup (at most 512 threads)");
  }
  // a filter step's unpack reads element (pad + j) ^ 1 for every slot j of
  // the launch, up to the widest row of the plan's threads per row (a batch
  // or a served launch may hold wider rows of other plans)
  for (int i = 0; i < n_ops; i++) {
    if (H.steps[i].kind != kFilter) continue;
    const size_t reach = (size_t)(P->step_pad[i] + pipeline_max_cols(P->threads) + 2) *
                         (P->dtype == PSEG_F32 ? sizeof(float) : sizeof(double));
    P->step_lds[i] = std::max(P->step_lds[i], reach);
    if (P->step_lds[i] > kMaxRowLds) {
      const size_t need = P->step_lds[i];
      delete P;
      return fail(PSEG_E_UNSUPPORTED, "op %d (%s): %zu bytes of LDS per row do not fit one "
                  "workgroup (at most %zu)", i, op_name(ops[i].kind), need, kMaxRowLds);
    }
  }
  // an IIR step's scan: one affine map (a 2 x 2 matrix and a vector, six
  // float64) per thread after the padded row
  for (int i = 0; i < n_ops; i++) {
    if (H.steps[i].kind != kIir) continue;
    P->step_lds[i] += (size_t)P->threads * 6 * sizeof(double);
    if (P->step_lds[i] > kMaxRowLds) {
      const size_t need = P->step_lds[i];
      delete P;
      return fail(PSEG_E_UNSUPPORTED, "op %d (%s): %zu bytes of LDS per row (the padded row in "
                  "float64) do not fit one workgroup (at most %zu)", i, op_name(ops[i].kind), need,
                  kMaxRowLds);
    }
  }
  // parts: a per-panel normalization after the first op starts one
  // (pipeline.nim:515-522)
  int start = 0;
  for (int i = 0; i < n_ops; i++) {
    if (H.steps[i].kind == kNormPanel && i > start) {
      P->part_a[P->n_parts] = start;
      P->part_b[P->n_parts] = i;
      P->n_parts++;
      start = i;
    }
  }
  P->part_a[P->n_parts] = start;
  P->part_b[P->n_parts] = n_ops;
  P->n_parts++;
  // the device copy
  char *base = nullptr;
  hipError_t e = hipMalloc((void **)&base, blob.bytes.size());
  if (e != hipSuccess) {
    delete P;
    return hip_fail(e, "hipMalloc of a plan");
  }
  for (int i = 0; i < n_ops; i++) patch_step(H.steps[i], base);
  memcpy(blob.bytes.data(), &H, sizeof H);
  e = hipMemcpy(base, blob.bytes.data(), blob.bytes.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(base);
    delete P;
    return hip_fail(e, "hipMemcpy of a plan");
  }
  P->dev = (DevPlan *)base;
  P->dev_base = base;
  P->generation = ++g_generation;
  *out = P;
  return PSEG_OK;
}

template <typename F>
pseg_status guarded(F f) {
  try {
    return f();
  } catch (const Invalid &e) {
    return fail(PSEG_E_INVALID, "%s", e.msg.c_str());
  } catch (const std::bad_alloc &) {
    return fail(PSEG_E_INVALID, "out of host memory");
  }
}

bool overlap(const void *a, const void *b, size_t bytes) {
  uintptr_t x = (uintptr_t)a, y = (uintptr_t)b;
  return x != y && x < y + bytes && y < x + bytes;
}

// The scratch of a per-panel normalization: row sums, row maxima, the
// histogram and the factor.
size_t panel_ws_bytes(int32_t rows) {
  size_t b = 0;
  b += ((size_t)rows * 8 + 255) / 256 * 256;       // row_sq
  b += ((size_t)rows * 8 + 255) / 256 * 256;       // row_max (2 x u32)
  b += 65536 * 4;                                   // histogram
  b += 256;                                         // factor
  return b;
}

}  // namespace

struct pseg_batch {
  int32_t n_items;
  int64_t rows, samples;
  int n_launches;
  // one launch per (dtype, threads per row, kernel form)
  int dtype[kMaxBatchClasses], threads[kMaxBatchClasses], full[kMaxBatchClasses];
  JobLaunch launch[kMaxBatchClasses];
  int resident[kMaxBatchClasses];  // workgroups of the class resident at once
  bool in_place;                   // an item runs in place (src == dst)
  char *dev;  // the job tables, then one 256-byte counter per launch
};

namespace {

// Relative cost of one row of a plan (FFT work dominates): heavier rows are
// handed out first, so the batch does not end on a long row started late.
double row_cost(const pseg_plan *P, int32_t cols) {
  double c = cols;
  for (int i = 0; i < P->n_steps; i++) {
    const int nf = P->step_nfft[i];
    if (nf > 0) c += 2.5 * nf * std::log2((double)nf);
    else if (P->host.steps[i].kind == kAgc) c += 8.0 * cols;
    else c += cols;
  }
  return c;
}

}  // namespace

namespace pseg {

pseg_status set_error(pseg_status st, const std::string &msg) {
  g_error = msg;
  return st;
}

pseg_status hip_error(hipError_t e, const char *what) { return hip_fail(e, what); }

}  // namespace pseg

// ==================================================================== C ABI

extern "C" {

size_t pseg_last_error(char *buf, size_t n) {
  if (buf && n) {
    size_t k = std::min(n - 1, g_error.size());
    memcpy(buf, g_error.data(), k);
    buf[k] = 0;
  }
  return g_error.size();
}

int32_t pseg_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

// Bytes per sample of a format the decode knows, else 0.
static int64_t format_bytes(int32_t format) {
  switch (format) {
    case PSEG_FMT_IBM32: case PSEG_FMT_INT32: case PSEG_FMT_IEEE32:
    case PSEG_FMT_SEGD_INT32: case PSEG_FMT_SEGD_IEEE32: return 4;
    case PSEG_FMT_INT16: return 2;
    case PSEG_FMT_INT8: return 1;
    case PSEG_FMT_SEGD_INT24: return 3;
    case PSEG_FMT_SEGD_IEEE64: return 8;
    default: return 0;
  }
}

pseg_status pseg_decode_samples(int32_t format, const void *raw, int64_t pitch, int64_t offset,
                                int32_t rows, int32_t cols, int32_t dtype, void *dst,
                                void *stream) {
  if (format_bytes(format) == 0)
    return fail(PSEG_E_INVALID, "unknown sample format %d (SEG-Y codes 1, 2, 3, 5, 8; SEG-D "
                "8036, 8038, 8058, 8080)", format);
  if (dtype != PSEG_F32 && dtype != PSEG_F64) return fail(PSEG_E_INVALID, "unknown dtype %d", dtype);
  if (rows < 0 || cols < 0) return fail(PSEG_E_INVALID, "rows and cols must be >= 0");
  if (rows == 0 || cols == 0) return PSEG_OK;
  if (!raw || !dst) return fail(PSEG_E_INVALID, "raw or dst is NULL");
  const int64_t bytes = format_bytes(format);
  if (offset < 0 || pitch < offset + (int64_t)cols * bytes)
    return fail(PSEG_E_INVALID, "%d samples of %lld bytes from byte %lld do not fit a trace of %lld "
                "bytes", cols, (long long)bytes, (long long)offset, (long long)pitch);
  if (bytes >= 4 && ((((uintptr_t)raw) | (uintptr_t)pitch | (uintptr_t)offset) & 3))
    return fail(PSEG_E_INVALID, "4- and 8-byte samples need raw, pitch and offset at multiples of 4");
  const size_t in_bytes = (size_t)(rows - 1) * (size_t)pitch + (size_t)offset + (size_t)cols * bytes;
  const size_t out_bytes = (size_t)rows * (size_t)cols * (dtype == PSEG_F32 ? 4 : 8);
  const uintptr_t a = (uintptr_t)raw, b = (uintptr_t)dst;
  if (a < b + out_bytes && b < a + in_bytes)
    return fail(PSEG_E_INVALID, "dst overlaps the raw traces");
  hipError_t e = launch_decode(dtype, format, raw, pitch, offset, rows, cols, dst, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "launch of the sample decode");
  return PSEG_OK;
}

pseg_status pseg_meta_flatten(int32_t rows, int32_t cols, int64_t first_sample,
                              const int64_t *num_samples, const uint8_t *status,
                              const int64_t *failed, int32_t n_failed,
                              int32_t for_correlation, int32_t *row_len) {
  if (rows < 0 || cols < 0) return fail(PSEG_E_INVALID, "rows and cols must be >= 0");
  if (first_sample < 0) return fail(PSEG_E_INVALID, "firstSample must be >= 0, got %lld", (long long)first_sample);
  if (rows > 0 && (!num_samples || !status || !row_len)) return fail(PSEG_E_INVALID, "NULL meta array");
  // initChunkMeta's checks of the failed list (core/types.nim:106-113)
  for (int32_t i = 0; i < n_failed; i++) {
    if (failed[i] < 0 || failed[i] >= rows)
      return fail(PSEG_E_INVALID, "failed row %lld is outside 0 ..< %d", (long long)failed[i], rows);
    if (i > 0 && failed[i] <= failed[i - 1])
      return fail(PSEG_E_INVALID, "failed rows must be ascending and unique: %lld is followed by %lld",
                  (long long)failed[i - 1], (long long)failed[i]);
  }
  int32_t f = 0;
  for (int32_t r = 0; r < rows; r++) {
    while (f < n_failed && failed[f] < r) f++;
    bool is_failed = f < n_failed && failed[f] == r;
    uint8_t s = status[r];
    bool ok = for_correlation == 2   ? true
              : for_correlation != 0 ? (!is_failed && s != 2)
                                     : (!is_failed && (s == 0 || s == 1));
    int64_t live = std::min<int64_t>(std::max<int64_t>(num_samples[r] - first_sample, 0), cols);
    row_len[r] = ok ? (int32_t)live : -1;
  }
  return PSEG_OK;
}

int64_t pseg_next_fast_len(int64_t n, int32_t even) { return next_fast(n, even != 0); }

pseg_status pseg_impulse_length(const pseg_filter_spec *spec, double dt_ms, int64_t *length) {
  if (!spec || !length) return fail(PSEG_E_INVALID, "NULL argument");
  return guarded([&] {
    *length = impulse_length(*spec, dt_ms);
    return PSEG_OK;
  });
}

double pseg_response(const pseg_filter_spec *spec, double f_hz, double dt_ms) {
  return spec ? response(*spec, f_hz, dt_ms) : NAN;
}

pseg_status pseg_plan_compile(const pseg_op *ops, int32_t n_ops, int32_t dtype, double dt_ms,
                              int32_t samples, int32_t record_samples, pseg_plan **out) {
  if (!out) return fail(PSEG_E_INVALID, "out is NULL");
  *out = nullptr;
  return guarded([&] {
    if (dtype == PSEG_F32) return compile_plan<float>(ops, n_ops, dt_ms, samples, record_samples, out);
    if (dtype == PSEG_F64) return compile_plan<double>(ops, n_ops, dt_ms, samples, record_samples, out);
    return fail(PSEG_E_INVALID, "dtype must be PSEG_F32 or PSEG_F64, got %d", dtype);
  });
}

pseg_status pseg_plan_destroy(pseg_plan *plan) {
  if (!plan) return PSEG_OK;
  if (plan->dev_base) (void)hipFree(plan->dev_base);
  delete plan;
  return PSEG_OK;
}

pseg_status pseg_plan_info_get(const pseg_plan *P, pseg_plan_info *o) {
  if (!P || !o) return fail(PSEG_E_INVALID, "NULL argument");
  memset(o, 0, sizeof *o);
  o->generation = P->generation;
  o->dtype = P->dtype;
  o->n_steps = P->n_steps;
  o->parts = P->n_parts;
  o->samples = P->samples;
  o->record_samples = P->record;
  o->margin = P->margin;
  o->dt_ms = P->dt_ms;
  size_t lds = 0;
  for (int i = 0; i < P->n_steps; i++) {
    o->step_kind[i] = P->step_kind[i];
    o->step_nfft[i] = P->step_nfft[i];
    o->step_pad[i] = P->step_pad[i];
    o->step_margin[i] = P->step_margin[i];
    o->step_reflect[i] = P->step_reflect[i];
    lds = std::max(lds, P->step_lds[i]);
  }
  o->lds_bytes = lds;
  o->threads_per_row = P->threads;
  o->amplitude_units = P->amplitude_units;
  o->n_calib = P->n_calib;
  return PSEG_OK;
}

pseg_status pseg_required_margin(const pseg_op *ops, int32_t n_ops, double dt_ms,
                                 int32_t *margin) {
  if (!margin) return fail(PSEG_E_INVALID, "margin is NULL");
  if (!std::isfinite(dt_ms) || dt_ms <= 0)
    return fail(PSEG_E_INVALID, "dtMs must be finite and positive, got %g", dt_ms);
  int64_t m = 0;
  for (int i = 0; i < n_ops; i++) {
    try {
      if (ops[i].kind == PSEG_OP_FILTER) {
        validate(ops[i].spec, dt_ms);
        m += impulse_length(ops[i].spec, dt_ms);
      } else if (ops[i].kind == PSEG_OP_AGC) {
        double h = std::round(ops[i].window_ms / (2 * dt_ms));
        if (!finite_positive(ops[i].window_ms))
          invalid("an AGC window must be finite and positive, got %g ms", ops[i].window_ms);
        if (h < 1) invalid("an AGC window of %g ms is under 3 samples at %g ms", ops[i].window_ms, dt_ms);
        m += (int64_t)h;
      } else if (ops[i].kind == PSEG_OP_IIR) {
        m += design_iir(ops[i].spec, dt_ms, ops[i].iir_phase).pad;
      }
    } catch (const Invalid &e) {
      return fail(PSEG_E_INVALID, "op %d (%s): %s", i, op_name(ops[i].kind), e.msg.c_str());
    }
  }
  *margin = (int32_t)m;
  return PSEG_OK;
}

void pseg_decode_window(int32_t shown_first, int32_t shown_last, int64_t margin,
                        int32_t trace_len, int32_t *first, int32_t *last) {
  int64_t m = std::min<int64_t>(std::max<int64_t>(margin, 0), std::max<int64_t>(trace_len, 0));
  if (first) *first = (int32_t)std::max<int64_t>(0, (int64_t)shown_first - m);
  if (last) *last = (int32_t)std::min<int64_t>((int64_t)trace_len - 1, (int64_t)shown_last + m);
}

pseg_status pseg_plan_workspace_size(const pseg_plan *P, int32_t rows, int32_t cols,
                                     size_t *bytes) {
  if (!P || !bytes) return fail(PSEG_E_INVALID, "NULL argument");
  (void)cols;
  bool panel = false;
  for (int i = 0; i < P->n_steps; i++) panel |= P->host.steps[i].kind == kNormPanel;
  *bytes = panel ? panel_ws_bytes(std::max(rows, 0)) : 0;
  return PSEG_OK;
}

pseg_status pseg_plan_apply(const pseg_plan *P, const void *src, void *dst, int32_t rows,
                            int32_t cols, const int32_t *row_len, void *ws, size_t ws_bytes,
                            void *stream) {
  return pseg_plan_apply_keys(P, src, dst, rows, cols, row_len, nullptr, 0, ws, ws_bytes, stream);
}

namespace {

const char *kNeedsCalib = "the plan has a calibration step: apply it with the rows' factors "
                          "(pseg_plan_apply_calib)";

pseg_status apply_plan(const pseg_plan *P, const void *src, void *dst, int32_t rows, int32_t cols,
                       const int32_t *row_len, const double *row_keys, int64_t first_sample,
                       const int32_t *cal_len, const double *cal_k, void *ws, size_t ws_bytes,
                       void *stream) {
  // checkApply (pipeline.nim:619-641): every check before any sample changes
  if (!P) return fail(PSEG_E_INVALID, "a plan must come from pseg_plan_compile");
  if (P->n_calib > 0 && (!cal_len || !cal_k) && rows > 0)
    return fail(PSEG_E_INVALID, "%s", kNeedsCalib);
  if (rows < 0 || cols < 0) return fail(PSEG_E_INVALID, "rows and cols must be >= 0");
  if (first_sample < 0) return fail(PSEG_E_INVALID, "firstSample must be >= 0, got %lld", (long long)first_sample);
  if (rows == 0) return PSEG_OK;
  if (!src || !dst || !row_len) return fail(PSEG_E_INVALID, "NULL panel or row_len");
  if (cols > P->samples)
    return fail(PSEG_E_INVALID, "the block has %d columns, the plan takes %d samples", cols, P->samples);
  size_t esz = P->dtype == PSEG_F32 ? 4 : 8;
  if (overlap(src, dst, (size_t)rows * (size_t)cols * esz))
    return fail(PSEG_E_INVALID, "src and dst overlap without being the same block");
  size_t need = 0;
  pseg_plan_workspace_size(P, rows, cols, &need);
  if (need > 0 && (!ws || ws_bytes < need))
    return fail(PSEG_E_WORKSPACE, "the plan needs a workspace of %zu bytes, got %zu", need,
                ws ? ws_bytes : (size_t)0);
  hipStream_t s = (hipStream_t)stream;
  char *w = (char *)ws;
  double *row_sq = (double *)w;
  uint32_t *row_max = (uint32_t *)(w + ((size_t)rows * 8 + 255) / 256 * 256);
  uint32_t *hist = (uint32_t *)((char *)row_max + ((size_t)rows * 8 + 255) / 256 * 256);
  double *factor = (double *)((char *)hist + 65536 * 4);
  for (int part = 0; part < P->n_parts; part++) {
    const int a = P->part_a[part], b = P->part_b[part];
    const void *in = part == 0 ? src : dst;
    JobLaunch L;
    memset(&L, 0, sizeof L);
    Job &J = L.single;
    J.plan = P->dev;
    J.src = in;
    J.dst = dst;
    J.num = row_len;
    J.keys = row_keys;
    if (P->n_calib > 0) {
      J.cal_len = cal_len;
      J.cal_k = cal_k;
    }
    J.col0 = first_sample;
    J.src_pitch = J.dst_pitch = cols;
    J.rows = rows;
    J.cols = cols;
    J.first = 0;
    J.crop0 = 0;
    J.crop1 = cols;
    J.step_first = a;
    J.step_last = b;
    const bool panel = b > a && P->host.steps[a].kind == kNormPanel;
    if (panel) {
      // panelFactorOf over the whole block as the previous parts left it
      hipError_t e = launch_panel_factor(P->dtype, in, rows, cols, row_len, P->norm_stat[a],
                                         P->norm_target[a], P->norm_pct[a], row_sq, row_max,
                                         hist, factor, s);
      if (e != hipSuccess) return hip_fail(e, "panel statistics");
      J.step_first = a + 1;
      J.has_factor = 1;
      J.factor = factor;
    }
    size_t lds = 1024;
    for (int i = J.step_first; i < J.step_last; i++) lds = std::max(lds, P->step_lds[i]);
    L.lds_bytes = lds;
    L.total = rows;
    L.max_cols = cols;
    for (int i = J.step_first; i < J.step_last; i++)
      L.full = std::max(L.full, step_form(P->host.steps[i].kind));
    hipError_t e = launch_pipeline(P->dtype, P->threads, L, s);
    if (e != hipSuccess) return hip_fail(e, "pipeline launch");
  }
  return PSEG_OK;
}

}  // namespace

pseg_status pseg_plan_apply_keys(const pseg_plan *P, const void *src, void *dst, int32_t rows,
                                 int32_t cols, const int32_t *row_len, const double *row_keys,
                                 int64_t first_sample, void *ws, size_t ws_bytes,
                                 void *stream) {
  return apply_plan(P, src, dst, rows, cols, row_len, row_keys, first_sample, nullptr, nullptr, ws,
                    ws_bytes, stream);
}

pseg_status pseg_plan_apply_calib(const pseg_plan *P, const void *src, void *dst, int32_t rows,
                                  int32_t cols, const int32_t *row_len, const double *row_keys,
                                  int64_t first_sample, const int32_t *calib_len,
                                  const double *calib_factors, void *ws, size_t ws_bytes,
                                  void *stream) {
  if (P && P->n_calib > 0 && rows > 0 && (!calib_len || !calib_factors))
    return fail(PSEG_E_INVALID, "NULL calib_len or calib_factors");
  return apply_plan(P, src, dst, rows, cols, row_len, row_keys, first_sample, calib_len,
                    calib_factors, ws, ws_bytes, stream);
}

pseg_status pseg_plan_apply_raw(const pseg_plan *P, int32_t format, const void *raw,
                                int64_t pitch, int64_t offset, void *dst, int32_t rows,
                                int32_t cols, const int32_t *row_len, const double *row_keys,
                                int64_t first_sample, void *stream) {
  if (!P) return fail(PSEG_E_INVALID, "a plan must come from pseg_plan_compile");
  if (format_bytes(format) == 0)
    return fail(PSEG_E_INVALID, "unknown sample format %d (SEG-Y codes 1, 2, 3, 5, 8; SEG-D "
                "8036, 8038, 8058, 8080)", format);
  if (rows < 0 || cols < 0) return fail(PSEG_E_INVALID, "rows and cols must be >= 0");
  if (first_sample < 0) return fail(PSEG_E_INVALID, "firstSample must be >= 0, got %lld", (long long)first_sample);
  for (int i = 0; i < P->n_steps; i++)
    if (P->host.steps[i].kind == kNormPanel)
      return fail(PSEG_E_UNSUPPORTED, "a plan with a per-panel normalization reads its panel "
                  "statistics from samples: decode first (pseg_decode_samples), then apply");
  if (P->n_calib > 0)
    return fail(PSEG_E_UNSUPPORTED, "a plan with a calibration step: decode first "
                "(pseg_decode_samples), then pseg_plan_apply_calib");
  if (rows == 0 || cols == 0) return PSEG_OK;
  if (!raw || !dst || !row_len) return fail(PSEG_E_INVALID, "NULL traces, panel or row_len");
  if (cols > P->samples)
    return fail(PSEG_E_INVALID, "the block has %d columns, the plan takes %d samples", cols, P->samples);
  const int64_t bytes = format_bytes(format);
  if (offset < 0 || pitch < offset + (int64_t)cols * bytes)
    return fail(PSEG_E_INVALID, "%d samples of %lld bytes from byte %lld do not fit a trace of %lld "
                "bytes", cols, (long long)bytes, (long long)offset, (long long)pitch);
  if (bytes >= 4 && ((((uintptr_t)raw) | (uintptr_t)pitch | (uintptr_t)offset) & 3))
    return fail(PSEG_E_INVALID, "4- and 8-byte samples need raw, pitch and offset at multiples of 4");
  const size_t esz = P->dtype == PSEG_F32 ? 4 : 8;
  const size_t in_bytes = (size_t)(rows - 1) * (size_t)pitch + (size_t)offset + (size_t)cols * bytes;
  const uintptr_t a = (uintptr_t)raw, b = (uintptr_t)dst;
  if (a < b + (size_t)rows * cols * esz && b < a + in_bytes)
    return fail(PSEG_E_INVALID, "dst overlaps the raw traces");
  JobLaunch L;
  memset(&L, 0, sizeof L);
  Job &J = L.single;
  J.plan = P->dev;
  J.src = raw;
  J.src_fmt = format;
  J.src_off = offset;
  J.src_pitch = pitch;
  J.dst = dst;
  J.num = row_len;
  J.keys = row_keys;
  J.col0 = first_sample;
  J.dst_pitch = cols;
  J.rows = rows;
  J.cols = cols;
  J.crop1 = cols;
  J.step_first = 0;
  J.step_last = P->n_steps;
  // the decoded row passes through LDS; the kernel form with every step
  size_t lds = std::max<size_t>(1024, (size_t)cols * esz);
  for (int i = 0; i < P->n_steps; i++) lds = std::max(lds, P->step_lds[i]);
  L.lds_bytes = lds;
  L.total = rows;
  L.max_cols = cols;
  L.full = std::max(1, P->full);
  hipError_t e = launch_pipeline(P->dtype, P->threads, L, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "pipeline launch");
  return PSEG_OK;
}

// --------------------------------------------------------------- spectrum

struct pseg_spectrum_plan {
  int dtype;
  int32_t window, nfft, bins, threads;
  size_t lds;
  DevSpec *dev;
  char *base;
};

}  // extern "C"

namespace {

template <typename T>
pseg_status spectrum_compile(double dt, int32_t window, double taper_fraction, int32_t nfft,
                             int32_t demean, pseg_spectrum_plan **out) {
  // initSpectrumPlan (fft/spectrum.nim:77-101)
  if (window < 1) return fail(PSEG_E_INVALID, "a spectrum window needs >= 1 sample, got %d", window);
  if (!(dt > 0 && dt < kInf)) return fail(PSEG_E_INVALID, "dtMs must be finite and > 0, got %g", dt);
  if (!(taper_fraction >= 0 && taper_fraction <= 0.5))
    return fail(PSEG_E_INVALID, "taperFraction must be in 0 .. 0.5, got %g", taper_fraction);
  int64_t n = nfft;
  if (n == 0) {
    n = 1;
    while (n < window) n *= 2;  // nextPowerOfTwo
  }
  if (n < window) return fail(PSEG_E_INVALID, "nfft %lld is shorter than the window (%d)", (long long)n, window);
  if (next_fast(n, false) != n) return fail(PSEG_E_INVALID, "nfft %lld is not 2·3·5-smooth", (long long)n);
  if (n > (1 << 20)) return fail(PSEG_E_UNSUPPORTED, "an FFT of %lld points is beyond this GPU path", (long long)n);
  int threads = 0;
  for (int t = 256; t <= 1024 && threads == 0; t *= 2)
    if (fft_fits(n % 2 ? n : n / 2, t, kFuseCorr, kBudgetCorr)) threads = t;
  const size_t lds = std::max(fft_lds_bytes(n, sizeof(T), kFuseCorr), (size_t)n * sizeof(T));
  // the FFT buffer and the workgroup's power sums (n / 2 + 1 float64)
  if (threads == 0 || (lds + 15) / 16 * 16 + (size_t)(n / 2 + 1) * 8 > kMaxRowLds)
    return fail(PSEG_E_UNSUPPORTED, "an FFT of %lld points does not fit one workgroup", (long long)n);
  Blob blob;
  blob.bytes.resize(sizeof(DevSpec));
  DevSpec H;
  memset(&H, 0, sizeof H);
  build_fft<T>(n, H.fft, blob, kFuseCorr);
  const double r = taper_fraction * (double)window;
  const int64_t ramp = (int64_t)std::floor(r + 0.5);  // round half away from zero (r >= 0)
  std::vector<T> taper(window);
  double sum = 0;
  for (int32_t i = 0; i < window; i++) {
    const int64_t d = std::min<int64_t>(i, window - 1 - i);
    double w = 1.0;
    if (d < ramp) {
      const double sn = std::sin(kPi / 2 * (double)(d + 1) / (double)(ramp + 1));
      w = sn * sn;
    }
    taper[i] = (T)w;
    sum += (double)taper[i];
  }
  set_off(H.taper, blob.add_vec(taper));
  H.taper_sum = sum;
  H.window = window;
  H.demean = demean != 0;
  char *base = nullptr;
  hipError_t e = hipMalloc((void **)&base, blob.bytes.size());
  if (e != hipSuccess) return hip_fail(e, "hipMalloc of a spectrum plan");
  patch_fft(H.fft, base);
  patch(H.taper, base);
  memcpy(blob.bytes.data(), &H, sizeof H);
  e = hipMemcpy(base, blob.bytes.data(), blob.bytes.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(base);
    return hip_fail(e, "hipMemcpy of a spectrum plan");
  }
  pseg_spectrum_plan *P = new pseg_spectrum_plan;
  P->dtype = sizeof(T) == 4 ? PSEG_F32 : PSEG_F64;
  P->window = window;
  P->nfft = (int32_t)n;
  P->bins = (int32_t)(n / 2 + 1);
  P->threads = threads;
  P->lds = lds;
  P->dev = (DevSpec *)base;
  P->base = base;
  *out = P;
  return PSEG_OK;
}

}  // namespace

extern "C" {

pseg_status pseg_spectrum_compile(int32_t dtype, double dt_ms, int32_t window,
                                  double taper_fraction, int32_t nfft, int32_t demean,
                                  pseg_spectrum_plan **out) {
  if (!out) return fail(PSEG_E_INVALID, "out is NULL");
  *out = nullptr;
  return guarded([&] {
    if (dtype == PSEG_F32) return spectrum_compile<float>(dt_ms, window, taper_fraction, nfft, demean, out);
    if (dtype == PSEG_F64) return spectrum_compile<double>(dt_ms, window, taper_fraction, nfft, demean, out);
    return fail(PSEG_E_INVALID, "unknown dtype %d", dtype);
  });
}

pseg_status pseg_spectrum_destroy(pseg_spectrum_plan *plan) {
  if (!plan) return PSEG_OK;
  (void)hipFree(plan->base);
  delete plan;
  return PSEG_OK;
}

pseg_status pseg_spectrum_info_get(const pseg_spectrum_plan *plan, int32_t *nfft, int32_t *bins) {
  if (!plan || !nfft || !bins) return fail(PSEG_E_INVALID, "NULL plan or output");
  *nfft = plan->nfft;
  *bins = plan->bins;
  return PSEG_OK;
}

pseg_status pseg_spectrum_workspace_size(const pseg_spectrum_plan *plan, int32_t rows,
                                         size_t *bytes) {
  if (!plan || !bytes) return fail(PSEG_E_INVALID, "NULL plan or output");
  if (rows < 0) return fail(PSEG_E_INVALID, "rows must be >= 0");
  *bytes = rows == 0 ? 0
                     : (size_t)spectrum_groups(rows, plan->bins) * ((size_t)plan->bins + 1) *
                           sizeof(double);
  return PSEG_OK;
}

pseg_status pseg_spectrum_average(const pseg_spectrum_plan *plan, const void *src, int32_t rows,
                                  int32_t cols, const int32_t *row_len, int32_t first_col,
                                  void *ws, size_t ws_bytes, double *amp, double *used,
                                  void *stream) {
  if (!plan) return fail(PSEG_E_INVALID, "a plan must come from pseg_spectrum_compile");
  if (rows < 0 || cols < 0) return fail(PSEG_E_INVALID, "rows and cols must be >= 0");
  if (first_col < 0) return fail(PSEG_E_INVALID, "firstCol must be >= 0, got %d", first_col);
  if (!amp) return fail(PSEG_E_INVALID, "NULL amp");
  if (rows > 0 && cols > 0 && (!src || !row_len)) return fail(PSEG_E_INVALID, "NULL panel or row_len");
  size_t need = 0;
  pseg_spectrum_workspace_size(plan, rows, &need);
  if (rows > 0 && cols > 0 && (!ws || ws_bytes < need))
    return fail(PSEG_E_WORKSPACE, "the workspace holds %zu bytes, the average of %d rows needs %zu",
                ws ? ws_bytes : (size_t)0, rows, need);
  if (((uintptr_t)ws | (uintptr_t)amp | (uintptr_t)used) % sizeof(double) != 0)
    return fail(PSEG_E_INVALID, "ws, amp and used must be 8-byte aligned");
  hipError_t e = launch_spectrum(plan->dtype, plan->threads, plan->dev, plan->lds, src, rows, cols,
                                 row_len, first_col, (double *)ws, amp, used, plan->bins,
                                 plan->nfft, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "spectrum launch");
  return PSEG_OK;
}

pseg_status pseg_spectrum_average_noalloc(const pseg_spectrum_plan *plan, const void *src,
                                          int32_t rows, int32_t cols, const int32_t *row_len,
                                          int32_t first_col, double *amp, double *used,
                                          void *stream) {
  if (!plan) return fail(PSEG_E_INVALID, "a plan must come from pseg_spectrum_compile");
  if (rows < 0 || cols < 0) return fail(PSEG_E_INVALID, "rows and cols must be >= 0");
  if (first_col < 0) return fail(PSEG_E_INVALID, "firstCol must be >= 0, got %d", first_col);
  if (!amp) return fail(PSEG_E_INVALID, "NULL amp");
  if (rows > 0 && cols > 0 && (!src || !row_len)) return fail(PSEG_E_INVALID, "NULL panel or row_len");
  if (((uintptr_t)amp | (uintptr_t)used) % sizeof(double) != 0)
    return fail(PSEG_E_INVALID, "amp and used must be 8-byte aligned");
  size_t need = 0;
  pseg_spectrum_workspace_size(plan, rows, &need);
  if (need > (size_t)kSpecPartials * sizeof(double))
    return fail(PSEG_E_INVALID,
                "the average of %d rows needs %zu bytes of partial sums, beyond the %zu of the "
                "allocation-free scratch", rows, need, (size_t)kSpecPartials * sizeof(double));
  hipError_t e = launch_spectrum_noalloc(plan->dtype, plan->threads, plan->dev, plan->lds, src,
                                         rows, cols, row_len, first_col, amp, used, plan->bins,
                                         plan->nfft, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "spectrum launch");
  return PSEG_OK;
}

struct pseg_spectrogram_plan {
  int dtype;
  int32_t window, hop, samples, frames, nfft, bins, threads;
  int32_t avg_threads;  // the averaged form: one wave per frame when its passes fit
  size_t lds;
  DevSpec *dev;
  char *base;
};

}  // extern "C"

namespace {

template <typename T>
pseg_status spectrogram_compile(int32_t window, int32_t hop, double dt, int32_t samples,
                                int32_t nfft, pseg_spectrogram_plan **out) {
  // initSpectrogramPlan (fft/spectrogram.nim)
  if (window < 2) return fail(PSEG_E_INVALID, "windowLen must be >= 2, got %d", window);
  if (hop < 1) return fail(PSEG_E_INVALID, "hop must be >= 1, got %d", hop);
  if (samples < 1) return fail(PSEG_E_INVALID, "samples must be >= 1, got %d", samples);
  if (!(dt > 0 && dt < kInf)) return fail(PSEG_E_INVALID, "dtMs must be finite and positive, got %g", dt);
  int64_t n = nfft;
  if (n == 0) {
    n = 1;
    while (n < window) n *= 2;
  }
  if (n < window) return fail(PSEG_E_INVALID, "nfft %lld is shorter than the window (%d)", (long long)n, window);
  if (next_fast(n, false) != n) return fail(PSEG_E_INVALID, "nfft %lld is not 2·3·5-smooth", (long long)n);
  if (n > (1 << 20)) return fail(PSEG_E_UNSUPPORTED, "an FFT of %lld points is beyond this GPU path", (long long)n);
  int threads = 0;
  for (int t = 256; t <= 1024 && threads == 0; t *= 2)
    if (fft_fits(n % 2 ? n : n / 2, t, kFuseCorr, kBudgetCorr)) threads = t;
  const size_t lds = std::max(fft_lds_bytes(n, sizeof(T), kFuseCorr), (size_t)n * sizeof(T));
  if (threads == 0 || lds > kMaxRowLds)
    return fail(PSEG_E_UNSUPPORTED, "an FFT of %lld points does not fit one workgroup", (long long)n);
  Blob blob;
  blob.bytes.resize(sizeof(DevSpec));
  DevSpec H;
  memset(&H, 0, sizeof H);
  build_fft<T>(n, H.fft, blob, kFuseCorr);
  std::vector<T> w(window);
  for (int32_t j = 0; j < window; j++) {
    const double sn = std::sin(kPi * (double)(j + 1) / (double)(window + 1));
    w[j] = (T)(sn * sn);
  }
  set_off(H.taper, blob.add_vec(w));
  H.window = window;
  char *base = nullptr;
  hipError_t e = hipMalloc((void **)&base, blob.bytes.size());
  if (e != hipSuccess) return hip_fail(e, "hipMalloc of a spectrogram plan");
  patch_fft(H.fft, base);
  patch(H.taper, base);
  memcpy(blob.bytes.data(), &H, sizeof H);
  e = hipMemcpy(base, blob.bytes.data(), blob.bytes.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(base);
    return hip_fail(e, "hipMemcpy of a spectrogram plan");
  }
  pseg_spectrogram_plan *P = new pseg_spectrogram_plan;
  P->dtype = sizeof(T) == 4 ? PSEG_F32 : PSEG_F64;
  P->window = window;
  P->hop = hop;
  P->samples = samples;
  P->frames = (samples - 1) / hop + 1;
  P->nfft = (int32_t)n;
  P->bins = (int32_t)(n / 2 + 1);
  P->threads = threads;
  // The averaged form runs many frames per CU at once: one wave per frame
  // when every pass fits 64 threads (a 512-point transform's fused passes
  // hold 16 butterflies, so 240 of 256 threads would idle in them).
  P->avg_threads = fft_fits(n % 2 ? n : n / 2, 64, kFuseCorr, kBudgetCorr) ? 64 : threads;
  P->lds = lds;
  P->dev = (DevSpec *)base;
  P->base = base;
  *out = P;
  return PSEG_OK;
}

}  // namespace

extern "C" {

pseg_status pseg_spectrogram_compile(int32_t dtype, int32_t window, int32_t hop, double dt_ms,
                                     int32_t samples, int32_t nfft, pseg_spectrogram_plan **out) {
  if (!out) return fail(PSEG_E_INVALID, "out is NULL");
  *out = nullptr;
  return guarded([&] {
    if (dtype == PSEG_F32) return spectrogram_compile<float>(window, hop, dt_ms, samples, nfft, out);
    if (dtype == PSEG_F64) return spectrogram_compile<double>(window, hop, dt_ms, samples, nfft, out);
    return fail(PSEG_E_INVALID, "unknown dtype %d", dtype);
  });
}

pseg_status pseg_spectrogram_destroy(pseg_spectrogram_plan *plan) {
  if (!plan) return PSEG_OK;
  (void)hipFree(plan->base);
  delete plan;
  return PSEG_OK;
}

pseg_status pseg_spectrogram_info_get(const pseg_spectrogram_plan *plan, int32_t *frames,
                                      int32_t *bins, int32_t *nfft) {
  if (!plan || !frames || !bins || !nfft) return fail(PSEG_E_INVALID, "NULL plan or output");
  *frames = plan->frames;
  *bins = plan->bins;
  *nfft = plan->nfft;
  return PSEG_OK;
}

namespace {
// The averaged spectrogram's workspace: status (rows int32), then the
// partials (groups·frames·bins float64) and the groups' row counts.
struct SgramWs {
  size_t status_off, parts_off, used_off, bytes;
  int32_t groups;
};
SgramWs sgram_ws(const pseg_spectrogram_plan *plan, int32_t rows) {
  SgramWs w;
  w.groups = sgram_groups(rows, plan->frames, plan->avg_threads);
  w.status_off = 0;
  w.parts_off = ((size_t)rows * sizeof(int32_t) + 255) / 256 * 256;
  w.used_off = w.parts_off + (size_t)w.groups * (size_t)plan->frames * (size_t)plan->bins * sizeof(double);
  w.bytes = rows == 0 ? 0 : w.used_off + (size_t)w.groups * sizeof(double);
  return w;
}
}  // namespace

pseg_status pseg_spectrogram_average_workspace_size(const pseg_spectrogram_plan *plan,
                                                    int32_t rows, size_t *bytes) {
  if (!plan || !bytes) return fail(PSEG_E_INVALID, "NULL plan or output");
  if (rows < 0) return fail(PSEG_E_INVALID, "rows must be >= 0");
  *bytes = sgram_ws(plan, rows).bytes;
  return PSEG_OK;
}

pseg_status pseg_spectrogram_average(const pseg_spectrogram_plan *plan, const void *src,
                                     int32_t rows, int32_t cols, const int32_t *row_len,
                                     void *ws, size_t ws_bytes, double *amp, double *used,
                                     void *stream) {
  if (!plan) return fail(PSEG_E_INVALID, "a plan must come from pseg_spectrogram_compile");
  if (rows < 0 || cols < 0) return fail(PSEG_E_INVALID, "rows and cols must be >= 0");
  if (cols > plan->samples)
    return fail(PSEG_E_INVALID, "rows of %d columns, the plan takes %d samples", cols, plan->samples);
  if (!amp) return fail(PSEG_E_INVALID, "NULL amp");
  if (((uintptr_t)ws | (uintptr_t)amp | (uintptr_t)used) % sizeof(double) != 0)
    return fail(PSEG_E_INVALID, "ws, amp and used must be 8-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = (int64_t)plan->frames * plan->bins;
  if (rows == 0 || cols == 0) {
    // nothing averaged: amp all zero, used 0
    hipError_t e = launch_zero_f64(amp, n, s);
    if (e == hipSuccess && used) e = launch_zero_f64(used, 1, s);
    if (e != hipSuccess) return hip_fail(e, "spectrogram launch");
    return PSEG_OK;
  }
  if (!src || !row_len) return fail(PSEG_E_INVALID, "NULL panel or row_len");
  const SgramWs w = sgram_ws(plan, rows);
  if (!ws || ws_bytes < w.bytes)
    return fail(PSEG_E_WORKSPACE, "the workspace holds %zu bytes, the average of %d rows needs %zu",
                ws ? ws_bytes : (size_t)0, rows, w.bytes);
  char *base = (char *)ws;
  hipError_t e = launch_spectrogram_average(
      plan->dtype, plan->avg_threads, plan->dev, plan->lds, plan->bins, src, rows, cols, row_len,
      plan->hop, plan->frames, w.groups, (int32_t *)(base + w.status_off),
      (double *)(base + w.parts_off), (double *)(base + w.used_off), amp, used, s);
  if (e != hipSuccess) return hip_fail(e, "spectrogram launch");
  return PSEG_OK;
}

pseg_status pseg_spectrogram_apply(const pseg_spectrogram_plan *plan, const void *x, int32_t len,
                                   double *amp, void *stream) {
  if (!plan) return fail(PSEG_E_INVALID, "a plan must come from pseg_spectrogram_compile");
  if (len < 0 || len > plan->samples)
    return fail(PSEG_E_INVALID, "a trace of %d samples, the plan takes 0 .. %d", len, plan->samples);
  if (!amp || (len > 0 && !x)) return fail(PSEG_E_INVALID, "NULL trace or amp");
  hipError_t e = launch_spectrogram(plan->dtype, plan->threads, plan->dev, plan->lds, x, len,
                                    plan->hop, plan->frames, amp, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "spectrogram launch");
  return PSEG_OK;
}

struct pseg_hilbert_plan {
  int dtype;
  int32_t samples, nfft, threads;
  size_t lds;
  DevFft *dev;
  char *base;
};

}  // extern "C"

namespace {

template <typename T>
pseg_status hilbert_compile(int32_t samples, int32_t nfft, pseg_hilbert_plan **out) {
  // initHilbertPlan (fft/hilbert.nim:44-53)
  if (samples < 1) return fail(PSEG_E_INVALID, "a Hilbert plan needs samples >= 1, got %d", samples);
  const int64_t n = nfft == 0 ? next_fast(2 * (int64_t)samples, true) : nfft;
  if (n < 2 * (int64_t)samples)
    return fail(PSEG_E_INVALID, "nfft %lld is below twice the samples (%lld): the transform would wrap",
                (long long)n, 2 * (long long)samples);
  if (next_fast(n, false) != n) return fail(PSEG_E_INVALID, "nfft %lld is not 2·3·5-smooth", (long long)n);
  if (n > (1 << 20)) return fail(PSEG_E_UNSUPPORTED, "an FFT of %lld points is beyond this GPU path", (long long)n);
  int threads = 0;
  for (int t = 256; t <= 1024 && threads == 0; t *= 2)
    if (fft_fits(n % 2 ? n : n / 2, t, kFuseCorr, kBudgetCorr)) threads = t;
  const size_t lds = std::max(fft_lds_bytes(n, sizeof(T), kFuseCorr), (size_t)n * sizeof(T));
  if (threads == 0 || lds > kMaxRowLds)
    return fail(PSEG_E_UNSUPPORTED, "an FFT of %lld points does not fit one workgroup", (long long)n);
  Blob blob;
  blob.bytes.resize(sizeof(DevFft));
  DevFft H;
  memset(&H, 0, sizeof H);
  build_fft<T>(n, H, blob, kFuseCorr);
  char *base = nullptr;
  hipError_t e = hipMalloc((void **)&base, blob.bytes.size());
  if (e != hipSuccess) return hip_fail(e, "hipMalloc of a Hilbert plan");
  patch_fft(H, base);
  memcpy(blob.bytes.data(), &H, sizeof H);
  e = hipMemcpy(base, blob.bytes.data(), blob.bytes.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(base);
    return hip_fail(e, "hipMemcpy of a Hilbert plan");
  }
  pseg_hilbert_plan *P = new pseg_hilbert_plan;
  P->dtype = sizeof(T) == 4 ? PSEG_F32 : PSEG_F64;
  P->samples = samples;
  P->nfft = (int32_t)n;
  P->threads = threads;
  P->lds = lds;
  P->dev = (DevFft *)base;
  P->base = base;
  *out = P;
  return PSEG_OK;
}

}  // namespace

extern "C" {

pseg_status pseg_hilbert_compile(int32_t dtype, int32_t samples, int32_t nfft,
                                 pseg_hilbert_plan **out) {
  if (!out) return fail(PSEG_E_INVALID, "out is NULL");
  *out = nullptr;
  return guarded([&] {
    if (dtype == PSEG_F32) return hilbert_compile<float>(samples, nfft, out);
    if (dtype == PSEG_F64) return hilbert_compile<double>(samples, nfft, out);
    return fail(PSEG_E_INVALID, "unknown dtype %d", dtype);
  });
}

pseg_status pseg_hilbert_destroy(pseg_hilbert_plan *plan) {
  if (!plan) return PSEG_OK;
  (void)hipFree(plan->base);
  delete plan;
  return PSEG_OK;
}

pseg_status pseg_hilbert_nfft(const pseg_hilbert_plan *plan, int32_t *nfft) {
  if (!plan || !nfft) return fail(PSEG_E_INVALID, "NULL plan or output");
  *nfft = plan->nfft;
  return PSEG_OK;
}

pseg_status pseg_hilbert_apply(const pseg_hilbert_plan *plan, const void *src, void *dst,
                               int32_t rows, int32_t cols, const int32_t *row_len,
                               int32_t envelope, void *stream) {
  if (!plan) return fail(PSEG_E_INVALID, "a plan must come from pseg_hilbert_compile");
  if (rows < 0 || cols < 0) return fail(PSEG_E_INVALID, "rows and cols must be >= 0");
  if (cols > plan->samples)
    return fail(PSEG_E_INVALID, "traces of %d columns, the plan takes %d samples", cols, plan->samples);
  if (rows > 0 && cols > 0 && (!src || !dst || !row_len))
    return fail(PSEG_E_INVALID, "NULL panel or row_len");
  const size_t esz = plan->dtype == PSEG_F32 ? 4 : 8;
  if (rows > 0 && cols > 0 && overlap(src, dst, (size_t)rows * (size_t)cols * esz))
    return fail(PSEG_E_INVALID, "src and dst overlap without being the same block");
  hipError_t e = launch_hilbert(plan->dtype, plan->threads, plan->dev, plan->lds, src, rows, cols,
                                row_len, envelope, dst, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "Hilbert launch");
  return PSEG_OK;
}

pseg_status pseg_hilbert_analytic(const pseg_hilbert_plan *plan, const void *src, int32_t rows,
                                  int32_t cols, const int32_t *row_len, double dt_ms,
                                  double floor_db, double *env, double *phase, double *hz,
                                  void *stream) {
  if (!plan) return fail(PSEG_E_INVALID, "a plan must come from pseg_hilbert_compile");
  if (rows < 0 || cols < 0) return fail(PSEG_E_INVALID, "rows and cols must be >= 0");
  if (cols > plan->samples)
    return fail(PSEG_E_INVALID, "traces of %d columns, the plan takes %d samples", cols, plan->samples);
  if (!(dt_ms > 0 && dt_ms < kInf)) return fail(PSEG_E_INVALID, "dtMs must be finite and > 0, got %g", dt_ms);
  if (std::isnan(floor_db)) return fail(PSEG_E_INVALID, "floorDb must not be NaN");
  if (rows > 0 && cols > 0 && (!src || !row_len || !env || !phase || !hz))
    return fail(PSEG_E_INVALID, "NULL panel, row_len, env, phase or hz");
  const double thr_scale = std::pow(10.0, floor_db / 20.0);
  hipError_t e = launch_analytic(plan->dtype, plan->threads, plan->dev, plan->lds, src, rows, cols,
                                 row_len, dt_ms, thr_scale, env, phase, hz, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "analytic launch");
  return PSEG_OK;
}

pseg_status pseg_spectrum_amplitude(const pseg_spectrum_plan *plan, const void *src, int32_t rows,
                                    int32_t cols, const int32_t *row_len, int32_t first_col,
                                    void *amp, void *phase, void *stream) {
  if (!plan) return fail(PSEG_E_INVALID, "a plan must come from pseg_spectrum_compile");
  if (rows < 0 || cols < 0) return fail(PSEG_E_INVALID, "rows and cols must be >= 0");
  if (first_col < 0) return fail(PSEG_E_INVALID, "firstCol must be >= 0, got %d", first_col);
  if (rows > 0 && cols > 0 && (!src || !row_len || !amp || !phase))
    return fail(PSEG_E_INVALID, "NULL panel, row_len, amp or phase");
  hipError_t e = launch_spectrum_rows(plan->dtype, plan->threads, plan->dev, plan->lds, src, rows,
                                      cols, row_len, first_col, amp, phase, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "spectrum launch");
  return PSEG_OK;
}

// ------------------------------------------------------------------ batch

pseg_status pseg_batch_create(const pseg_batch_item *items, int32_t n, pseg_batch **out) {
  if (!out) return fail(PSEG_E_INVALID, "out is NULL");
  *out = nullptr;
  if (n < 0) return fail(PSEG_E_INVALID, "n_items must be >= 0, got %d", n);
  if (n > 0 && !items) return fail(PSEG_E_INVALID, "items is NULL");
  // every check before anything runs (pipeline.nim:619-641, per item)
  for (int32_t i = 0; i < n; i++) {
    const pseg_batch_item &it = items[i];
    const pseg_plan *P = it.plan;
    if (!P) return fail(PSEG_E_INVALID, "item %d: a plan must come from pseg_plan_compile", i);
    if (it.rows < 0 || it.cols < 0)
      return fail(PSEG_E_INVALID, "item %d: rows and cols must be >= 0", i);
    if (it.rows == 0) continue;
    if (!it.src || !it.dst || !it.row_len)
      return fail(PSEG_E_INVALID, "item %d: NULL panel or row_len", i);
    if (it.first_sample < 0)
      return fail(PSEG_E_INVALID, "item %d: firstSample must be >= 0, got %lld", i,
                  (long long)it.first_sample);
    if (it.cols > P->samples)
      return fail(PSEG_E_INVALID, "item %d: the block has %d columns, the plan takes %d samples",
                  i, it.cols, P->samples);
    const size_t esz = P->dtype == PSEG_F32 ? 4 : 8;
    if (overlap(it.src, it.dst, (size_t)it.rows * (size_t)it.cols * esz))
      return fail(PSEG_E_INVALID, "item %d: src and dst overlap without being the same block", i);
    if (P->n_parts != 1)
      return fail(PSEG_E_UNSUPPORTED, "item %d: a plan with a per-panel normalization runs on its "
                  "own (pseg_plan_apply)", i);
    if (P->n_calib > 0)
      return fail(PSEG_E_UNSUPPORTED, "item %d: a plan with a calibration step runs on its own "
                  "(pseg_plan_apply_calib)", i);
  }
  // no item writes where another reads or writes
  for (int32_t i = 0; i < n; i++) {
    const pseg_batch_item &a = items[i];
    if (a.rows == 0 || a.cols == 0) continue;
    const size_t ab = (size_t)a.rows * a.cols * (a.plan->dtype == PSEG_F32 ? 4 : 8);
    for (int32_t j = 0; j < n; j++) {
      const pseg_batch_item &b = items[j];
      if (j == i || b.rows == 0 || b.cols == 0) continue;
      const size_t bb = (size_t)b.rows * b.cols * (b.plan->dtype == PSEG_F32 ? 4 : 8);
      const uintptr_t d = (uintptr_t)a.dst;
      const uintptr_t s2 = (uintptr_t)b.src, d2 = (uintptr_t)b.dst;
      if ((d < d2 + bb && d2 < d + ab) || (d < s2 + bb && s2 < d + ab))
        return fail(PSEG_E_INVALID, "items %d and %d: item %d's dst overlaps item %d's samples",
                    std::min(i, j), std::max(i, j), i, j);
    }
  }
  pseg_batch *B = new pseg_batch;
  memset(B, 0, sizeof *B);
  B->n_items = n;
  // one launch per (dtype, threads per row)
  std::vector<std::vector<int>> groups;
  for (int32_t i = 0; i < n; i++) {
    const pseg_batch_item &it = items[i];
    if (it.rows == 0) continue;
    B->rows += it.rows;
    B->samples += (int64_t)it.rows * it.cols;
    int g = 0;
    for (; g < B->n_launches; g++)
      if (B->dtype[g] == it.plan->dtype && B->threads[g] == it.plan->threads &&
          B->full[g] == it.plan->full)
        break;
    if (g == B->n_launches) {
      B->dtype[g] = it.plan->dtype;
      B->threads[g] = it.plan->threads;
      B->full[g] = it.plan->full;
      B->n_launches++;
      groups.emplace_back();
    }
    groups[g].push_back(i);
  }
  std::vector<std::vector<Job>> tables(B->n_launches);
  size_t bytes = 0;
  for (int g = 0; g < B->n_launches; g++) {
    std::vector<int> &ids = groups[g];
    std::stable_sort(ids.begin(), ids.end(), [&](int a, int b) {
      return row_cost(items[a].plan, items[a].cols) > row_cost(items[b].plan, items[b].cols);
    });
    JobLaunch &L = B->launch[g];
    size_t lds = 1024;
    int32_t item0 = 0;
    for (int id : ids) {
      const pseg_batch_item &it = items[id];
      const pseg_plan *P = it.plan;
      Job J;
      memset(&J, 0, sizeof J);
      J.plan = P->dev;
      J.src = it.src;
      J.dst = it.dst;
      J.num = it.row_len;
      J.keys = it.row_keys;
      J.col0 = it.first_sample;
      J.src_pitch = J.dst_pitch = it.cols;
      J.rows = it.rows;
      J.cols = it.cols;
      J.crop1 = it.cols;
      J.item0 = item0;
      J.step_first = 0;
      J.step_last = P->n_steps;
      tables[g].push_back(J);
      if ((int64_t)item0 + it.rows > INT32_MAX) {
        delete B;
        return fail(PSEG_E_INVALID, "a batch holds at most %d rows per class", INT32_MAX);
      }
      item0 += it.rows;
      for (int s = 0; s < P->n_steps; s++) lds = std::max(lds, P->step_lds[s]);
      L.max_cols = std::max(L.max_cols, it.cols);
    }
    L.n_jobs = (int32_t)ids.size();
    L.total = item0;
    L.lds_bytes = lds;
    L.full = B->full[g];
    int resident = 0;
    hipError_t e = pipeline_resident_blocks(B->dtype[g], B->threads[g], L.max_cols, lds, L.full,
                                            &resident);
    if (e != hipSuccess) {
      delete B;
      return hip_fail(e, "occupancy of the pipeline kernel");
    }
    L.grid = std::max(1, std::min(L.total, resident));
    B->resident[g] = std::max(1, resident);
    bytes += (tables[g].size() * sizeof(Job) + 255) / 256 * 256;
  }
  bytes += 256 * (size_t)std::max(B->n_launches, 1);
  hipError_t e = hipMalloc((void **)&B->dev, bytes);
  if (e != hipSuccess) {
    delete B;
    return hip_fail(e, "hipMalloc of a batch");
  }
  size_t off = 0;
  for (int g = 0; g < B->n_launches; g++) {
    const size_t tb = tables[g].size() * sizeof(Job);
    e = hipMemcpy(B->dev + off, tables[g].data(), tb, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      (void)hipFree(B->dev);
      delete B;
      return hip_fail(e, "hipMemcpy of a batch");
    }
    B->launch[g].jobs = (const Job *)(B->dev + off);
    off += (tb + 255) / 256 * 256;
  }
  for (int g = 0; g < B->n_launches; g++) B->launch[g].counter = (uint32_t *)(B->dev + off + 256 * g);
  B->in_place = false;
  for (int32_t i = 0; i < n; i++) B->in_place |= items[i].src == items[i].dst;
  *out = B;
  return PSEG_OK;
}

// The launch shape of each class by timing (no counterpart in the
// reference: its CPU pool has no launch shape).  Every candidate gives each
// row the same arithmetic (a row's bits depend on its plan and samples only),
// so the choice changes the time of a batch, never its results.
pseg_status pseg_batch_tune(pseg_batch *B, int32_t repeats, void *stream, pseg_batch_tuning *out) {
  if (!B) return fail(PSEG_E_INVALID, "a batch must come from pseg_batch_create");
  if (repeats < 1) return fail(PSEG_E_INVALID, "repeats must be >= 1, got %d", repeats);
  if (B->in_place)
    return fail(PSEG_E_INVALID, "tuning runs the batch many times: an item in place would be "
                "processed again each time (tune a copy src -> dst)");
  hipStream_t s = (hipStream_t)stream;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e != hipSuccess) return hip_fail(e, "hipEventCreate");
  pseg_batch_tuning t;
  memset(&t, 0, sizeof t);
  t.launches = B->n_launches;
  pseg_status st = PSEG_OK;
  for (int g = 0; g < B->n_launches && st == PSEG_OK; g++) {
    JobLaunch &L = B->launch[g];
    uint32_t *counter = L.counter;
    const int res = B->resident[g];
    // persistent grids of 1/2, 3/4 and all of the resident workgroups, and
    // the direct launch (one workgroup per row, the dispatcher's order)
    const int grids[4] = {std::max(1, res / 2), std::max(1, 3 * res / 4), res, L.total};
    float best = 0;
    int best_i = -1;
    for (int c = 0; c < 4 && st == PSEG_OK; c++) {
      const bool direct = c == 3;
      if (!direct && grids[c] >= L.total && c < 2) continue;
      L.grid = std::max(1, std::min(grids[c], L.total));
      L.counter = direct ? nullptr : counter;
      std::vector<float> ms;
      for (int k = 0; k <= repeats && st == PSEG_OK; k++) {  // k = 0 warms up
        hipError_t le = hipEventRecord(e0, s);
        if (le == hipSuccess) le = launch_pipeline(B->dtype[g], B->threads[g], L, s);
        if (le == hipSuccess) le = hipEventRecord(e1, s);
        if (le == hipSuccess) le = hipEventSynchronize(e1);
        float m = 0;
        if (le == hipSuccess) le = hipEventElapsedTime(&m, e0, e1);
        if (le != hipSuccess) st = hip_fail(le, "batch tuning launch");
        else if (k > 0) ms.push_back(m);
      }
      if (st != PSEG_OK) break;
      std::sort(ms.begin(), ms.end());
      const float med = ms[ms.size() / 2];
      t.candidate_ms[g][c] = med;
      if (best_i < 0 || med < best) {
        best = med;
        best_i = c;
      }
    }
    if (st != PSEG_OK) {
      L.grid = std::max(1, std::min(L.total, res));
      L.counter = counter;
      break;
    }
    L.grid = std::max(1, std::min(grids[best_i], L.total));
    L.counter = best_i == 3 ? nullptr : counter;
    t.grid[g] = L.grid;
    t.direct[g] = best_i == 3;
    t.best_ms[g] = best;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st == PSEG_OK && out) *out = t;
  return st;
}

pseg_status pseg_batch_launch(const pseg_batch *B, void *stream) {
  if (!B) return fail(PSEG_E_INVALID, "a batch must come from pseg_batch_create");
  for (int g = 0; g < B->n_launches; g++) {
    hipError_t e = launch_pipeline(B->dtype[g], B->threads[g], B->launch[g], (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "batch launch");
  }
  return PSEG_OK;
}

pseg_status pseg_batch_info_get(const pseg_batch *B, pseg_batch_info *o) {
  if (!B || !o) return fail(PSEG_E_INVALID, "NULL argument");
  memset(o, 0, sizeof *o);
  o->items = B->n_items;
  o->launches = B->n_launches;
  o->rows = B->rows;
  o->samples = B->samples;
  for (int g = 0; g < B->n_launches; g++) {
    o->threads[g] = B->threads[g];
    o->grid[g] = B->launch[g].counter ? B->launch[g].grid : 0;
  }
  return PSEG_OK;
}

pseg_status pseg_batch_destroy(pseg_batch *B) {
  if (!B) return PSEG_OK;
  if (B->dev) (void)hipFree(B->dev);
  delete B;
  return PSEG_OK;
}

// ------------------------------------------------------------ correlation

pseg_status pseg_corr_compile(int32_t dtype, const double *pilot, int32_t plen, double pdt,
                              double pt0, double dt, int32_t record, int32_t out_samples,
                              int32_t mode, double eps_percent, int32_t normalize,
                              pseg_corr_plan **out) {
  if (!out) return fail(PSEG_E_INVALID, "out is NULL");
  *out = nullptr;
  if (dtype != PSEG_F32 && dtype != PSEG_F64)
    return fail(PSEG_E_INVALID, "dtype must be PSEG_F32 or PSEG_F64, got %d", dtype);
  // Pilot.validate (vibroseis/pilot.nim:89-106)
  if (!pilot || plen < 2) return fail(PSEG_E_INVALID, "a pilot needs at least two samples, got %d", plen);
  if (!(pdt > 0 && pdt < kInf)) return fail(PSEG_E_INVALID, "a pilot's dtMs must be finite and positive, got %g", pdt);
  if (!std::isfinite(pt0)) return fail(PSEG_E_INVALID, "a pilot's t0Ms must be finite, got %g", pt0);
  double energy = 0;
  for (int32_t i = 0; i < plen; i++) {
    if (!std::isfinite(pilot[i])) return fail(PSEG_E_INVALID, "pilot sample %d is not finite", i);
    energy += pilot[i] * pilot[i];
  }
  if (energy == 0) return fail(PSEG_E_INVALID, "the pilot is all zeros");
  // initCorrelationPlan (vibroseis/correlate.nim:94-160)
  if (!(dt > 0 && dt < kInf)) return fail(PSEG_E_INVALID, "dtMs must be finite and positive, got %g", dt);
  if (!(std::fabs(pdt - dt) <= 1e-9 * dt))
    return fail(PSEG_E_INVALID, "the pilot is sampled at %g ms, the records at %g ms: resample the pilot first", pdt, dt);
  if (record < 1) return fail(PSEG_E_INVALID, "recordSamples must be >= 1, got %d", record);
  int64_t outN = out_samples == 0 ? (int64_t)record - plen : out_samples;
  if (outN < 1)
    return fail(PSEG_E_INVALID, "the pilot (%d samples) is not shorter than the record (%d): no listen time; pass outSamples", plen, record);
  if (out_samples < 0 || outN > record)
    return fail(PSEG_E_INVALID, "outSamples must be in 0 .. recordSamples, got %d", out_samples);
  if (mode == 1 && !(eps_percent > 0 && eps_percent < kInf))
    return fail(PSEG_E_INVALID, "epsPercent must be finite and positive, got %g", eps_percent);
  if (mode != 0 && mode != 1) return fail(PSEG_E_INVALID, "unknown correlation mode %d", mode);
  double sft = pt0 / dt;
  if (std::fabs(sft) >= (double)((int64_t)record + plen))
    return fail(PSEG_E_INVALID, "the pilot starts %g ms from the time break: no lag of a %d-sample record reaches the output", pt0, record);
  int64_t whole = (int64_t)std::round(sft);
  double frac = sft - (double)whole;
  int64_t n = next_fast((int64_t)record + plen + 64 + std::llabs(whole), true);
  size_t esz = dtype == PSEG_F32 ? 4 : 8;
  size_t lds = fft_lds_bytes(n, esz, kFuseCorr);
  // 512 threads per record when every pass fits them (two records then
  // share a CU when their LDS allows), else 1024.
  const int threads = fft_fits(n / 2, 512, kFuseCorr, kBudgetCorr) ? 512 : 1024;
  if (!fft_fits(n / 2, threads, kFuseCorr, kBudgetCorr))
    return fail(PSEG_E_UNSUPPORTED, "a correlation FFT of %lld points is too long for the "
                "in-place transform of one workgroup", (long long)n);
  if (lds > kMaxRowLds)
    return fail(PSEG_E_UNSUPPORTED, "a correlation FFT of %lld points does not fit one workgroup's LDS", (long long)n);
  // The pilot's spectrum in float64.
  HostFft hf(n);
  std::vector<cd> x(n, cd(0, 0));
  for (int32_t j = 0; j < plen; j++) x[j] = cd(pilot[j], 0);
  std::vector<cd> P = hf.forward(x);
  const int64_t bins = n / 2 + 1;
  double peak_pow = 0;
  for (int64_t k = 0; k < bins; k++) peak_pow = std::max(peak_pow, std::norm(P[k]));
  double eps = mode == 1 ? eps_percent / 100 * peak_pow : 0.0;
  std::vector<double> w(bins);
  double peak = 0;
  for (int64_t k = 0; k < bins; k++) {
    double p2 = P[k].real() * P[k].real() + P[k].imag() * P[k].imag();
    w[k] = mode == 1 ? 1 / (p2 + eps) : 1.0;
    double g = p2 * w[k];
    peak += (k == 0 || (n % 2 == 0 && k == bins - 1)) ? g : 2 * g;
  }
  peak /= (double)n;
  double scale = normalize ? 1 / peak : 1.0;
  Blob blob;
  blob.bytes.resize(sizeof(DevCorr));
  DevCorr C;
  memset(&C, 0, sizeof C);
  try {
    if (dtype == PSEG_F32) build_fft<float>(n, C.fft, blob, kFuseCorr);
    else build_fft<double>(n, C.fft, blob, kFuseCorr);
  } catch (const Invalid &e) {
    return fail(PSEG_E_INVALID, "%s", e.msg.c_str());
  }
  // conj(P) w scale exp(+2 pi i k frac / n)
  std::vector<cd> mul(bins);
  // the device folds the split's 1/2 and the inverse's 1/n into the multiplier
  const double fold = 0.5 / (double)n;
  for (int64_t k = 0; k < bins; k++) {
    double ws = w[k] * scale * fold;
    double ar = P[k].real() * ws, ai = -P[k].imag() * ws;
    double th = 2 * kPi * (double)k * frac / (double)n;
    double er = std::cos(th), ei = std::sin(th);
    mul[k] = cd(ar * er - ai * ei, ar * ei + ai * er);
  }
  // the inverse takes Nyquist as real (only its real part enters the coefficients)
  if (dtype == PSEG_F32) set_off(C.coef, blob.add_vec(lin_coef_cplx<float>(mul, n)));
  else set_off(C.coef, blob.add_vec(lin_coef_cplx<double>(mul, n)));
  C.shift = (int32_t)whole;
  C.pilot = plen;
  C.record = record;
  C.out = (int32_t)outN;
  char *base = nullptr;
  hipError_t e = hipMalloc((void **)&base, blob.bytes.size());
  if (e != hipSuccess) return hip_fail(e, "hipMalloc of a correlation plan");
  patch_fft(C.fft, base);
  patch(C.coef, base);
  memcpy(blob.bytes.data(), &C, sizeof C);
  e = hipMemcpy(base, blob.bytes.data(), blob.bytes.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(base);
    return hip_fail(e, "hipMemcpy of a correlation plan");
  }
  pseg_corr_plan *cp = new pseg_corr_plan;
  cp->dtype = dtype;
  cp->dt_ms = dt;
  cp->scale = scale;
  cp->nfft = (int32_t)n;
  cp->shift = (int32_t)whole;
  cp->out = (int32_t)outN;
  cp->pilot = plen;
  cp->record = record;
  cp->lds = lds;
  cp->threads = threads;
  cp->dev = (DevCorr *)base;
  *out = cp;
  return PSEG_OK;
}

pseg_status pseg_corr_destroy(pseg_corr_plan *plan) {
  if (!plan) return PSEG_OK;
  if (plan->dev) (void)hipFree(plan->dev);
  delete plan;
  return PSEG_OK;
}

pseg_status pseg_corr_info_get(const pseg_corr_plan *c, pseg_corr_info *o) {
  if (!c || !o) return fail(PSEG_E_INVALID, "NULL argument");
  memset(o, 0, sizeof *o);
  o->dtype = c->dtype;
  o->nfft = c->nfft;
  o->shift = c->shift;
  o->out_samples = c->out;
  o->pilot_samples = c->pilot;
  o->record_samples = c->record;
  o->dt_ms = c->dt_ms;
  o->scale = c->scale;
  o->lds_bytes = c->lds;
  o->threads_per_record = c->threads;
  return PSEG_OK;
}

pseg_status pseg_corr_apply(const pseg_corr_plan *c, const void *src, int32_t rows,
                            int32_t src_cols, const int32_t *row_len, void *dst,
                            int32_t dst_cols, void *stream) {
  // checkCorrelate (vibroseis/correlate.nim:328-347)
  if (!c) return fail(PSEG_E_INVALID, "a plan must come from pseg_corr_compile");
  if (rows < 0 || src_cols < 0) return fail(PSEG_E_INVALID, "rows and cols must be >= 0");
  if (src_cols > c->record)
    return fail(PSEG_E_INVALID, "the records are %d samples, the plan takes %d", src_cols, c->record);
  if (dst_cols != c->out)
    return fail(PSEG_E_INVALID, "dst must be %d samples wide, got %d", c->out, dst_cols);
  if (rows == 0) return PSEG_OK;
  if (!src || !dst || !row_len) return fail(PSEG_E_INVALID, "NULL panel or row_len");
  hipError_t e = launch_correlate(c->dtype, c->threads, c->dev, c->lds, src, rows, src_cols, row_len, dst,
                                  (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "correlation launch");
  return PSEG_OK;
}

pseg_status pseg_linear_sweep(double f1, double f2, double length_ms, double dt_ms, double ts,
                              double te, double phase_deg, double amp, double *x, int32_t n) {
  // validate (vibroseis/sweep.nim:132-183) for a linear law
  if (!(dt_ms > 0 && dt_ms < kInf)) return fail(PSEG_E_INVALID, "dtMs must be finite and positive, got %g", dt_ms);
  if (!(length_ms >= dt_ms && length_ms < kInf))
    return fail(PSEG_E_INVALID, "lengthMs must be at least one sample interval, got %g", length_ms);
  double steps = length_ms / dt_ms;
  if (std::fabs(steps - std::round(steps)) > 1e-6)
    return fail(PSEG_E_INVALID, "lengthMs = %g must be a whole number of samples of %g ms", length_ms, dt_ms);
  if (!std::isfinite(amp)) return fail(PSEG_E_INVALID, "the amplitude must be finite, got %g", amp);
  if (!std::isfinite(phase_deg)) return fail(PSEG_E_INVALID, "startPhaseDeg must be finite, got %g", phase_deg);
  double nyq = 500 / dt_ms, limit = 0.8 * nyq;
  for (double f : {f1, f2}) {
    if (!(f > 0 && f < kInf)) return fail(PSEG_E_INVALID, "a sweep frequency must be > 0 Hz, got %g", f);
    if (f > nyq) return fail(PSEG_E_INVALID, "%g Hz is above Nyquist (%g Hz): it would alias", f, nyq);
    if (f > limit * (1 + 1e-12))
      return fail(PSEG_E_INVALID, "%g Hz is above the anti-alias limit %g Hz", f, limit);
  }
  double tsl = std::isnan(ts) ? 0 : ts, tel = std::isnan(te) ? 0 : te;
  if (!(tsl >= 0 && tsl < kInf) || !(tel >= 0 && tel < kInf))
    return fail(PSEG_E_INVALID, "a taper length must be >= 0 and finite");
  if (tsl + tel > length_ms * (1 + 1e-9))
    return fail(PSEG_E_INVALID, "the sweep tapers overlap: %g + %g ms is longer than the %g ms sweep", tsl, tel, length_ms);
  // samplesInLength (core/types.nim:147-166)
  double q = length_ms / dt_ms;
  double r = std::round(q);
  int64_t len = std::fabs(q - r) <= 1e-6 ? (int64_t)r : (int64_t)std::floor(q);
  if (!x || n != len) return fail(PSEG_E_INVALID, "the sweep has %lld samples, the buffer %d", (long long)len, n);
  // phasesInto (slLinear) and sweepInto (vibroseis/sweep.nim:292-342)
  const double dt = dt_ms / 1000, T = length_ms / 1000, rate = (f2 - f1) / T;
  const double p0 = phase_deg * kPi / 180;
  for (int32_t j = 0; j < n; j++) {
    double t = (double)j * dt;
    x[j] = amp * std::sin(p0 + 2 * kPi * (f1 * t + 0.5 * rate * t * t));
  }
  // applyTaper with tpCos2 (vibroseis/taper.nim:125-179)
  auto wgt = [](double v) {
    v = std::min(1.0, std::max(0.0, v));
    return 0.5 - 0.5 * std::cos(kPi * v);
  };
  const double sweep_ms = (double)n * dt_ms;
  if (tsl > 0)
    for (int32_t j = 0; j < n && (double)j * dt_ms < tsl; j++) x[j] = x[j] * wgt((double)j * dt_ms / tsl);
  if (tel > 0)
    for (int32_t j = n - 1; j >= 0 && sweep_ms - (double)j * dt_ms < tel; j--)
      x[j] = x[j] * wgt((sweep_ms - (double)j * dt_ms) / tel);
  return PSEG_OK;
}

}  // extern "C"

// ============================================================ first breaks

struct pseg_firstbreak_plan {
  int dtype;
  int64_t samples;
  DevFirstBreak host;
  DevFirstBreak *dev;
};

namespace {

bool fb_positive(double x) { return x > 0 && x < kInf; }

// samplesIn (qc/firstbreak.nim:189-192): a window in samples, rounded as a
// float first and held to 1 .. samples + 1.
int64_t fb_samples_in(double ms, double dt, int64_t samples) {
  return (int64_t)std::min(std::max(std::round(ms / dt), 1.0), (double)(samples + 1));
}

// initFirstBreakPlan (qc/firstbreak.nim:195-251): the checks and messages,
// the windows in samples, the band-pass (minimum phase, trend removed only
// when the whole trace is read), and what a pick reads about its window.
pseg_status fb_compile(int dtype, const pseg_fb_params &p, double dt, int64_t samples,
                       pseg_firstbreak_plan **out) {
  if (!fb_positive(dt)) return fail(PSEG_E_INVALID, "dtMs must be finite and positive");
  if (samples < 1) return fail(PSEG_E_INVALID, "a first-break plan needs samples >= 1");
  if (samples > INT32_MAX) return fail(PSEG_E_INVALID, "samples must fit an int32");
  if (!(fb_positive(p.window_ms) && fb_positive(p.lta_ms) && fb_positive(p.snr_window_ms)))
    return fail(PSEG_E_INVALID, "first-break windows must be finite and positive");
  if (p.lta_ms < p.window_ms) return fail(PSEG_E_INVALID, "ltaMs must be >= windowMs");
  if (!(p.stabilization >= 0 && p.stabilization < kInf))
    return fail(PSEG_E_INVALID, "stabilization must be finite and >= 0");
  if (!(p.low_db < p.high_db) || p.low_db == -kInf || p.high_db == kInf)
    return fail(PSEG_E_INVALID, "need finite lowDb < highDb");
  if (p.search_end_ms < p.search_start_ms) return fail(PSEG_E_INVALID, "searchEndMs < searchStartMs");
  if (p.search_start_ms == kInf || p.search_end_ms == -kInf)
    return fail(PSEG_E_INVALID, "the search window is empty");
  if (p.method < 0 || p.method > 2) return fail(PSEG_E_INVALID, "unknown pick method %d", p.method);
  if (!std::isnan(p.guide_velocity)) {
    if (!fb_positive(p.guide_velocity))
      return fail(PSEG_E_INVALID, "guideVelocity must be finite and positive (or NaN)");
    if (!(p.before_ms >= 0 && p.after_ms >= 0 && p.before_ms < kInf && p.after_ms < kInf))
      return fail(PSEG_E_INVALID, "beforeMs and afterMs must be finite and >= 0");
    if (!std::isfinite(p.guide_t0_ms)) return fail(PSEG_E_INVALID, "guideT0Ms must be finite");
  }
  if (!(p.low_cut_hz >= 0 && p.high_cut_hz > p.low_cut_hz))
    return fail(PSEG_E_INVALID, "need 0 <= lowCutHz < highCutHz");
  DevFirstBreak D;
  memset(&D, 0, sizeof D);
  D.method = p.method;
  D.samples = samples;
  D.dt = dt;
  D.w = fb_samples_in(p.window_ms, dt, samples);
  D.lta = fb_samples_in(p.lta_ms, dt, samples);
  D.snr = fb_samples_in(p.snr_window_ms, dt, samples);
  D.stabilization = p.stabilization;
  D.search_start = p.search_start_ms;
  D.search_end = p.search_end_ms;
  D.guide_v = p.guide_velocity;
  D.guide_t0 = p.guide_t0_ms;
  D.before_ms = p.before_ms;
  D.after_ms = p.after_ms;
  D.low_db = p.low_db;
  D.high_db = p.high_db;
  D.filtered = p.low_cut_hz > 0 || p.high_cut_hz < kInf;
  D.whole = p.method == 2 || (std::isnan(p.guide_velocity) && std::isnan(p.search_start_ms) &&
                              std::isnan(p.search_end_ms));
  int64_t pad = 0;
  if (D.filtered) {
    pseg_filter_spec spec;
    memset(&spec, 0, sizeof spec);
    spec.kind = 1;
    spec.f3 = spec.f4 = kInf;
    spec.low_hz = p.low_cut_hz;
    spec.low_slope = p.slope_db_per_octave;
    spec.high_hz = p.high_cut_hz;
    spec.high_slope = p.slope_db_per_octave;
    pseg_status st = guarded([&] {
      const IirDesign d = design_iir(spec, dt, 1);
      if ((int)d.secs.size() > kFbMaxSections)
        return fail(PSEG_E_UNSUPPORTED, "a band-pass of %d sections", (int)d.secs.size());
      D.n_sections = (int32_t)d.secs.size();
      for (size_t i = 0; i < d.secs.size(); i++) {
        const Sos &s = d.secs[i];
        const double v[6] = {s.b0, s.b1, s.b2, s.a1, s.a2, s.g};
        for (int k = 0; k < 6; k++) D.sec[i][k] = v[k];
      }
      pad = d.pad;
      D.dc_gain = d.dc_gain;
      D.dc_slope = d.dc_slope;
      return PSEG_OK;
    });
    if (st != PSEG_OK) return st;
    D.trend = D.whole ? 2 : 0;
    D.reflect = std::fabs(D.dc_gain) >= 0.5 || D.trend != 2;
    D.lead = D.reflect ? pad : 0;
  }
  const int64_t need = p.method == 0 ? D.w + D.lta : D.w;
  D.before = std::max(D.snr + D.w / 2, need) + pad;
  D.after = std::max(D.w, D.snr);
  DevFirstBreak *dev = nullptr;
  hipError_t e = hipMalloc((void **)&dev, sizeof D);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc of a first-break plan");
  e = hipMemcpy(dev, &D, sizeof D, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(dev);
    return hip_fail(e, "hipMemcpy of a first-break plan");
  }
  pseg_firstbreak_plan *P = new pseg_firstbreak_plan;
  P->dtype = dtype;
  P->samples = samples;
  P->host = D;
  P->dev = dev;
  *out = P;
  return PSEG_OK;
}

// The scratch of a pick: prefix energies and the attribute (cols + 1
// doubles a row each), the filter's padded trace (lead + cols + 1 doubles)
// and the filtered copy (cols samples), each [index][rows].
struct FbWs {
  size_t p, a, buf, copy, total;
};

FbWs fb_ws(const pseg_firstbreak_plan *P, int32_t rows, int32_t cols) {
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t r = (size_t)rows, c = (size_t)cols + 1;
  FbWs w;
  w.p = 0;
  w.a = up(r * c * 8);
  w.buf = w.a + up(r * c * 8);
  const size_t nbuf = P->host.filtered ? r * ((size_t)P->host.lead + c) * 8 : 0;
  w.copy = w.buf + up(nbuf);
  const size_t ncopy = P->host.filtered ? r * (size_t)cols * (P->dtype == PSEG_F32 ? 4 : 8) : 0;
  w.total = w.copy + up(ncopy);
  return w;
}

}  // namespace

extern "C" {

pseg_status pseg_firstbreak_compile(int32_t dtype, const pseg_fb_params *params, double dt_ms,
                                    int64_t samples, pseg_firstbreak_plan **out) {
  if (!out) return fail(PSEG_E_INVALID, "out is NULL");
  *out = nullptr;
  if (!params) return fail(PSEG_E_INVALID, "params is NULL");
  if (dtype != PSEG_F32 && dtype != PSEG_F64) return fail(PSEG_E_INVALID, "unknown dtype %d", dtype);
  return fb_compile(dtype, *params, dt_ms, samples, out);
}

pseg_status pseg_firstbreak_destroy(pseg_firstbreak_plan *plan) {
  if (!plan) return PSEG_OK;
  (void)hipFree(plan->dev);
  delete plan;
  return PSEG_OK;
}

pseg_status pseg_firstbreak_required_margin(const pseg_firstbreak_plan *plan, int64_t *out) {
  if (!plan || !out) return fail(PSEG_E_INVALID, "NULL plan or output");
  *out = plan->host.whole ? plan->samples : plan->host.before;
  return PSEG_OK;
}

pseg_status pseg_firstbreak_workspace_size(const pseg_firstbreak_plan *plan, int32_t rows,
                                           int32_t cols, size_t *out) {
  if (!plan || !out) return fail(PSEG_E_INVALID, "NULL plan or output");
  if (rows < 0 || cols < 0) return fail(PSEG_E_INVALID, "rows and cols must be >= 0");
  *out = fb_ws(plan, rows, cols).total;
  return PSEG_OK;
}

pseg_status pseg_firstbreak_pick(const pseg_firstbreak_plan *plan, const void *src, int32_t rows,
                                 int32_t cols, const int32_t *row_len, const double *keys,
                                 int64_t first_sample, void *ws, size_t ws_bytes, double *time_ms,
                                 int64_t *col, double *snr_db, double *conf, void *stream) {
  if (!plan) return fail(PSEG_E_INVALID, "a plan must come from pseg_firstbreak_compile");
  if (rows < 0 || cols < 0) return fail(PSEG_E_INVALID, "rows and cols must be >= 0");
  if (cols > plan->samples)
    return fail(PSEG_E_INVALID, "the block has %d columns, the plan takes %lld", cols,
                (long long)plan->samples);
  if (first_sample < 0) return fail(PSEG_E_INVALID, "firstSample must be >= 0");
  if (rows == 0) return PSEG_OK;
  if (!row_len || !time_ms || !col || !snr_db || !conf || (cols > 0 && !src))
    return fail(PSEG_E_INVALID, "NULL panel, row_len or pick outputs");
  const FbWs w = fb_ws(plan, rows, cols);
  if (!ws || ws_bytes < w.total || ((uintptr_t)ws & 255))
    return fail(PSEG_E_INVALID, "ws must be 256-byte aligned and hold %zu bytes, got %zu",
                w.total, ws_bytes);
  char *b = (char *)ws;
  hipError_t e = launch_firstbreak(plan->dtype, plan->dev, src, rows, cols, row_len, keys,
                                   first_sample, (double *)(b + w.p), (double *)(b + w.a),
                                   (double *)(b + w.buf), b + w.copy, time_ms, col, snr_db, conf,
                                   (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "first-break launch");
  return PSEG_OK;
}

}  // extern "C"

// ================================================================ fold maps

extern "C" {

pseg_status pseg_fold_add(const pseg_binning *binning, const double *bounds, int32_t n_bounds,
                          int32_t sectors, double sector_start_deg, int32_t offset_range,
                          const double *keys, const int32_t *row_len, int32_t rows,
                          int32_t *fold, int32_t *offset_fold, int32_t *azimuth_fold,
                          float *offset_min, float *offset_max, uint64_t *counters,
                          void *stream) {
  if (!binning) return fail(PSEG_E_INVALID, "binning is NULL");
  const pseg_binning &b = *binning;
  if (b.kind < PSEG_BIN_CDP || b.kind > PSEG_BIN_INLINE_CROSSLINE)
    return fail(PSEG_E_INVALID, "unknown binning kind %d", b.kind);
  if (b.inlines < 1 || b.crosslines < 1 || b.inlines > INT64_MAX / b.crosslines)
    return fail(PSEG_E_INVALID, "a binning needs positive inline and crossline counts");
  if ((b.kind == PSEG_BIN_INLINE_CROSSLINE || b.kind == PSEG_BIN_CDP) &&
      !(b.step0 > 0 && b.step0 < kInf && (b.kind == PSEG_BIN_CDP || (b.step1 > 0 && b.step1 < kInf))))
    return fail(PSEG_E_INVALID, "binning steps must be finite and positive");
  if ((b.kind == PSEG_BIN_LINE || b.kind == PSEG_BIN_GRID) &&
      !(b.size0 > 0 && b.size0 < kInf && (b.kind == PSEG_BIN_LINE || (b.size1 > 0 && b.size1 < kInf))))
    return fail(PSEG_E_INVALID, "bin sizes must be finite and positive");
  if (b.kind != PSEG_BIN_GRID && b.kind != PSEG_BIN_INLINE_CROSSLINE && b.inlines != 1)
    return fail(PSEG_E_INVALID, "a 2D binning has one inline");
  if (rows < 0) return fail(PSEG_E_INVALID, "rows must be >= 0");
  if (n_bounds < 0 || n_bounds == 1) return fail(PSEG_E_INVALID, "offset bounds need 0 or >= 2 edges");
  if (sectors < 0) return fail(PSEG_E_INVALID, "azimuthSectors must be >= 0");
  if (!std::isfinite(sector_start_deg)) return fail(PSEG_E_INVALID, "sectorStartDeg must be finite");
  if (!fold || !counters || (n_bounds > 1 && (!bounds || !offset_fold)) ||
      (sectors > 0 && !azimuth_fold) || (offset_range && (!offset_min || !offset_max)))
    return fail(PSEG_E_INVALID, "NULL fold, counters, or an array its options need");
  if (rows > 0 && (!keys || !row_len)) return fail(PSEG_E_INVALID, "NULL keys or row_len");
  hipError_t e = launch_fold_add(b, bounds, n_bounds, sectors, sector_start_deg, offset_range ? 1 : 0,
                                 keys, row_len, rows, fold, offset_fold, azimuth_fold, offset_min,
                                 offset_max, (unsigned long long *)counters, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "fold launch");
  return PSEG_OK;
}

pseg_status pseg_fold_clear(int64_t bins, int32_t classes, int32_t sectors, int32_t offset_range,
                            int32_t *fold, int32_t *offset_fold, int32_t *azimuth_fold,
                            float *offset_min, float *offset_max, uint64_t *counters,
                            void *stream) {
  if (bins < 1 || classes < 0 || sectors < 0) return fail(PSEG_E_INVALID, "bins >= 1, classes and sectors >= 0");
  if (!fold || !counters || (classes > 0 && !offset_fold) || (sectors > 0 && !azimuth_fold) ||
      (offset_range && (!offset_min || !offset_max)))
    return fail(PSEG_E_INVALID, "NULL fold, counters, or an array its options need");
  hipError_t e = launch_fold_clear(bins, classes, sectors, offset_range ? 1 : 0, fold, offset_fold,
                                   azimuth_fold, offset_min, offset_max,
                                   (unsigned long long *)counters, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "fold clear launch");
  return PSEG_OK;
}

pseg_status pseg_fold_difference(const int32_t *pre, const int32_t *actual, int32_t *out,
                                 int64_t bins, void *stream) {
  if (bins < 0) return fail(PSEG_E_INVALID, "bins must be >= 0");
  if (bins > 0 && (!pre || !actual || !out)) return fail(PSEG_E_INVALID, "NULL fold array");
  hipError_t e = launch_fold_diff(pre, actual, out, bins, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "fold difference launch");
  return PSEG_OK;
}

}  // extern "C"

// =========================================================== host staging

namespace {

// Device memory pseg_plan_apply_host stages a host block through: one
// region per device, per thread, grown to the largest block seen.
struct Staging {
  int device;
  char *base;
  size_t bytes;
};
thread_local std::vector<Staging> t_staging;

char *staging_for(size_t bytes, hipError_t &e) {
  int dev = 0;
  e = hipGetDevice(&dev);
  if (e != hipSuccess) return nullptr;
  for (Staging &st : t_staging) {
    if (st.device != dev) continue;
    if (st.bytes >= bytes) return st.base;
    (void)hipFree(st.base);
    st.base = nullptr;
    st.bytes = 0;
    e = hipMalloc((void **)&st.base, bytes);
    if (e != hipSuccess) return nullptr;
    st.bytes = bytes;
    return st.base;
  }
  Staging st{dev, nullptr, 0};
  e = hipMalloc((void **)&st.base, bytes);
  if (e != hipSuccess) return nullptr;
  st.bytes = bytes;
  t_staging.push_back(st);
  return st.base;
}

size_t up256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

extern "C" {

pseg_status pseg_plan_apply_host(const pseg_plan *P, const void *src, void *dst, int32_t rows,
                                 int32_t cols, const int32_t *row_len, const double *row_keys,
                                 int64_t first_sample, void *stream) {
  // checkApply (pipeline.nim:619-641) before any copy
  if (!P) return fail(PSEG_E_INVALID, "a plan must come from pseg_plan_compile");
  if (P->n_calib > 0 && rows > 0) return fail(PSEG_E_INVALID, "%s", kNeedsCalib);
  if (rows < 0 || cols < 0) return fail(PSEG_E_INVALID, "rows and cols must be >= 0");
  if (first_sample < 0) return fail(PSEG_E_INVALID, "firstSample must be >= 0, got %lld", (long long)first_sample);
  if (rows == 0) return PSEG_OK;
  if (!src || !dst || !row_len) return fail(PSEG_E_INVALID, "NULL panel or row_len");
  if (cols > P->samples)
    return fail(PSEG_E_INVALID, "the block has %d columns, the plan takes %d samples", cols, P->samples);
  const size_t esz = P->dtype == PSEG_F32 ? 4 : 8;
  const size_t panel = (size_t)rows * (size_t)cols * esz;
  if (overlap(src, dst, panel))
    return fail(PSEG_E_INVALID, "src and dst overlap without being the same block");
  size_t need = 0;
  pseg_plan_workspace_size(P, rows, cols, &need);
  const size_t o_len = up256(panel), o_keys = o_len + up256((size_t)rows * 4);
  const size_t o_ws = o_keys + (row_keys ? up256((size_t)rows * 16) : 0);
  const size_t total = o_ws + up256(std::max<size_t>(need, 1));
  hipError_t e;
  char *base = staging_for(total, e);
  if (!base) return hip_fail(e, "hipMalloc of the host staging block");
  hipStream_t s = (hipStream_t)stream;
  if (panel > 0 && (e = hipMemcpyAsync(base, src, panel, hipMemcpyHostToDevice, s)) != hipSuccess)
    return hip_fail(e, "copy of the block to the device");
  if ((e = hipMemcpyAsync(base + o_len, row_len, (size_t)rows * 4, hipMemcpyHostToDevice, s)) != hipSuccess)
    return hip_fail(e, "copy of row_len to the device");
  if (row_keys &&
      (e = hipMemcpyAsync(base + o_keys, row_keys, (size_t)rows * 16, hipMemcpyHostToDevice, s)) != hipSuccess)
    return hip_fail(e, "copy of the row keys to the device");
  pseg_status st = pseg_plan_apply_keys(P, base, base, rows, cols, (const int32_t *)(base + o_len),
                                        row_keys ? (const double *)(base + o_keys) : nullptr,
                                        first_sample, need ? base + o_ws : nullptr, need, stream);
  if (st != PSEG_OK) return st;
  if (panel > 0 && (e = hipMemcpyAsync(dst, base, panel, hipMemcpyDeviceToHost, s)) != hipSuccess)
    return hip_fail(e, "copy of the block to the host");
  if ((e = hipStreamSynchronize(s)) != hipSuccess) return hip_fail(e, "host staging");
  return PSEG_OK;
}

pseg_status pseg_host_staging_release(void) {
  for (Staging &st : t_staging) {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(st.device);
    (void)hipFree(st.base);
    (void)hipSetDevice(cur);
  }
  t_staging.clear();
  return PSEG_OK;
}

}  // extern "C"

// -------------------------------------------------------------------- f-k

struct pseg_fk_plan {
  int dtype;
  int32_t traces, samples, n_time, n_k, n_freq, threads;
  double dt_ms;
  size_t lds_time, lds_space;
  DevFk *dev;
  char *base;
};

namespace {

// tukey (fft/fk.nim:70-80): 1 in the middle, 0.5 - 0.5 cos(pi (d + 0.5) /
// ramp) over ramp = round(fraction n) samples at each end.
template <typename T>
std::vector<T> fk_tukey(int32_t n, double fraction) {
  std::vector<T> w(n);
  const int64_t ramp = fraction <= 0 ? 0 : (int64_t)std::floor(fraction * (double)n + 0.5);
  for (int32_t i = 0; i < n; i++) {
    const int64_t d = std::min<int64_t>(i, n - 1 - i);
    const double v = (ramp <= 0 || d >= ramp) ? 1.0
                     : 0.5 - 0.5 * std::cos(kPi * ((double)d + 0.5) / (double)ramp);
    w[i] = (T)v;
  }
  return w;
}

template <typename T>
pseg_status fk_compile(double dt, int32_t traces, int32_t samples, double taper,
                       pseg_fk_plan **out) {
  // initFkPlan (fft/fk.nim:82-106)
  if (traces < 4) return fail(PSEG_E_INVALID, "an f-k gather needs ≥ 4 live rows, got %d", traces);
  if (samples < 8) return fail(PSEG_E_INVALID, "an f-k window needs ≥ 8 samples, got %d", samples);
  if (!(dt > 0 && dt < kInf)) return fail(PSEG_E_INVALID, "dtMs must be finite and > 0, got %g", dt);
  if (!(taper >= 0 && taper <= 0.5))
    return fail(PSEG_E_INVALID, "taperFraction must be in 0 .. 0.5, got %g", taper);
  const int64_t nt = next_fast(samples, true), nk = next_fast(2 * (int64_t)traces, true);
  int threads = 0;
  for (int t = 256; t <= 1024 && threads == 0; t *= 2)
    if (fft_fits(nt / 2, t, kFuseCorr, kBudgetCorr) && fft_fits(nk / 2, t, kFuseCorr, kBudgetCorr))
      threads = t;
  const size_t lds_time = std::max(fft_lds_bytes(nt, sizeof(T), kFuseCorr), (size_t)nt * sizeof(T));
  const size_t space = (std::max(fft_lds_bytes(nk, sizeof(T), kFuseCorr), (size_t)nk * sizeof(T)) +
                        15) / 16 * 16;
  const size_t keep = ((size_t)(nk / 2 + 1) * 2 * sizeof(T) + 15) / 16 * 16;
  // the column's imaginary parts beside the transform, so the spatial pass
  // reads each (strided) spectrum value once, when that fits
  const size_t im = (size_t)traces * sizeof(T);
  const bool keep_im = space + keep + im <= kMaxRowLds;
  const size_t lds_space = space + keep + (keep_im ? im : 0);
  if (threads == 0 || lds_time > kMaxRowLds || lds_space > kMaxRowLds)
    return fail(PSEG_E_UNSUPPORTED, "an f-k transform of %lld x %lld points does not fit one "
                "workgroup", (long long)nt, (long long)nk);
  Blob blob;
  blob.bytes.resize(sizeof(DevFk));
  DevFk H;
  memset(&H, 0, sizeof H);
  build_fft<T>(nt, H.time, blob, kFuseCorr);
  build_fft<T>(nk, H.space, blob, kFuseCorr);
  set_off(H.time_w, blob.add_vec(fk_tukey<T>(samples, taper)));
  set_off(H.trace_w, blob.add_vec(fk_tukey<T>(traces, taper)));
  H.traces = traces;
  H.samples = samples;
  H.n_freq = (int32_t)(nt / 2 + 1);
  H.n_k = (int32_t)nk;
  H.space_lds = (int64_t)space;
  H.im_lds = keep_im ? (int64_t)(space + keep) : 0;
  char *base = nullptr;
  hipError_t e = hipMalloc((void **)&base, blob.bytes.size());
  if (e != hipSuccess) return hip_fail(e, "hipMalloc of an f-k plan");
  patch_fft(H.time, base);
  patch_fft(H.space, base);
  patch(H.time_w, base);
  patch(H.trace_w, base);
  memcpy(blob.bytes.data(), &H, sizeof H);
  e = hipMemcpy(base, blob.bytes.data(), blob.bytes.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(base);
    return hip_fail(e, "hipMemcpy of an f-k plan");
  }
  pseg_fk_plan *P = new pseg_fk_plan;
  P->dtype = sizeof(T) == 4 ? PSEG_F32 : PSEG_F64;
  P->traces = traces;
  P->samples = samples;
  P->n_time = (int32_t)nt;
  P->n_k = (int32_t)nk;
  P->n_freq = H.n_freq;
  P->threads = threads;
  P->dt_ms = dt;
  P->lds_time = lds_time;
  P->lds_space = lds_space;
  P->dev = (DevFk *)base;
  P->base = base;
  *out = P;
  return PSEG_OK;
}

size_t fk_spec_bytes(const pseg_fk_plan *P) {
  const size_t esz = P->dtype == PSEG_F32 ? 4 : 8;
  return ((size_t)P->traces * P->n_freq * 2 * esz + 255) / 256 * 256;
}

}  // namespace

extern "C" {

pseg_status pseg_fk_compile(int32_t dtype, double dt_ms, int32_t traces, int32_t samples,
                            double taper_fraction, pseg_fk_plan **out) {
  if (!out) return fail(PSEG_E_INVALID, "out is NULL");
  *out = nullptr;
  return guarded([&] {
    if (dtype == PSEG_F32) return fk_compile<float>(dt_ms, traces, samples, taper_fraction, out);
    if (dtype == PSEG_F64) return fk_compile<double>(dt_ms, traces, samples, taper_fraction, out);
    return fail(PSEG_E_INVALID, "unknown dtype %d", dtype);
  });
}

pseg_status pseg_fk_destroy(pseg_fk_plan *plan) {
  if (!plan) return PSEG_OK;
  (void)hipFree(plan->base);
  delete plan;
  return PSEG_OK;
}

pseg_status pseg_fk_axes(const pseg_fk_plan *P, int32_t *n_freq, int32_t *n_k, double *df_hz,
                         double *dk_per_trace) {
  if (!P) return fail(PSEG_E_INVALID, "a plan must come from pseg_fk_compile");
  if (n_freq) *n_freq = P->n_freq;
  if (n_k) *n_k = P->n_k;
  if (df_hz) *df_hz = 1000 / ((double)P->n_time * P->dt_ms);
  if (dk_per_trace) *dk_per_trace = 1 / (double)P->n_k;
  return PSEG_OK;
}

pseg_status pseg_fk_workspace_size(const pseg_fk_plan *P, size_t *bytes) {
  if (!P || !bytes) return fail(PSEG_E_INVALID, "NULL argument");
  *bytes = fk_spec_bytes(P) + 256;
  return PSEG_OK;
}

pseg_status pseg_fk_amplitude(const pseg_fk_plan *P, const void *src, int32_t rows, int32_t cols,
                              const int32_t *row_len, const int32_t *live_rows, int32_t first_col,
                              double *amp, void *ws, size_t ws_bytes, void *stream) {
  if (!P) return fail(PSEG_E_INVALID, "a plan must come from pseg_fk_compile");
  if (first_col < 0) return fail(PSEG_E_INVALID, "firstCol must be ≥ 0, got %d", first_col);
  if (rows < 0 || cols < 0) return fail(PSEG_E_INVALID, "rows and cols must be >= 0");
  if (!src || !row_len || !live_rows || !amp) return fail(PSEG_E_INVALID, "NULL argument");
  size_t need = 0;
  pseg_fk_workspace_size(P, &need);
  if (!ws || ws_bytes < need)
    return fail(PSEG_E_WORKSPACE, "the plan needs a workspace of %zu bytes, got %zu", need,
                ws ? ws_bytes : (size_t)0);
  char *w = (char *)ws;
  hipError_t e = launch_fk(P->dtype, P->threads, P->dev, P->lds_time, P->lds_space, P->traces,
                           P->n_freq, src, cols, row_len, live_rows, first_col, w,
                           (uint32_t *)(w + fk_spec_bytes(P)), amp, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(e, "f-k launch");
  return PSEG_OK;
}

}  // extern "C"

// ------------------------------------------------------- applied response

extern "C" {

// |Σ x[j] exp(-2πi f j dt)| at each frequency (responseModuli,
// pipeline.nim:774-838): a unit phasor advanced by the rotation exp(-2πi f
// dt) each sample, then turned onto the exact angle -2π f (j+1) dt by the
// first-order correction for the rounding of that step, so no drift builds up
// over long traces; one sine and cosine per frequency.
pseg_status pseg_response_moduli(const double *x, int32_t n, const double *hz, int32_t m,
                                 double dt_s, double *gain) {
  if (n < 0 || m < 0) return fail(PSEG_E_INVALID, "n and m must be >= 0");
  if ((n > 0 && !x) || (m > 0 && (!hz || !gain))) return fail(PSEG_E_INVALID, "NULL argument");
  for (int32_t i = 0; i < m; i++) {
    const double a = -2.0 * kPi * hz[i];
    const double w = a * dt_s, cw = std::cos(w), sw = std::sin(w);
    double c = 1, s = 0, at = 0, re = 0, im = 0;
    for (int32_t j = 0; j < n; j++) {
      re += x[j] * c;
      im += x[j] * s;
      const double next = (a * (double)(j + 1)) * dt_s;  // the exact angle of sample j + 1
      const double step = next - at;                      // this step, as rounded
      const double e = step - w;                          // its difference from w
      const double rc = cw - sw * e, rs = sw + cw * e;    // rotation by w + e
      const double nc = c * rc - s * rs, ns = c * rs + s * rc;
      const double d = next - (at + step);                // what the step left over
      c = nc - ns * d;
      s = ns + nc * d;
      at = next;
    }
    gain[i] = std::hypot(re, im);
  }
  return PSEG_OK;
}

// appliedResponseInto (pipeline.nim:840-880): the flow compiled with every
// FFT and IIR filter's trend removal off, applied on the current device to a
// unit impulse at sample samples/2 of one live row (t0Ms 0, offset 0, no file
// scales), and the moduli of that output's DFT at hz.
pseg_status pseg_applied_response(const pseg_op *ops, int32_t n_ops, int32_t dtype, double dt_ms,
                                  int32_t samples, const double *hz, int32_t m, double *gain) {
  if (n_ops < 0 || n_ops > PSEG_MAX_OPS || (n_ops > 0 && !ops))
    return fail(PSEG_E_INVALID, "a flow has 0 .. %d ops", PSEG_MAX_OPS);
  if (m < 0 || (m > 0 && (!hz || !gain))) return fail(PSEG_E_INVALID, "NULL hz or gain");
  std::vector<pseg_op> copy(ops, ops + n_ops);
  for (pseg_op &o : copy)
    if (o.kind == PSEG_OP_FILTER || o.kind == PSEG_OP_IIR) o.trend = 0;
  pseg_plan *P = nullptr;
  pseg_status st = pseg_plan_compile(copy.data(), n_ops, dtype, dt_ms, samples, 0, &P);
  if (st != PSEG_OK) return st;
  const size_t esz = dtype == PSEG_F32 ? 4 : 8;
  std::vector<char> host((size_t)samples * esz, 0);
  if (dtype == PSEG_F32) ((float *)host.data())[samples / 2] = 1;
  else ((double *)host.data())[samples / 2] = 1;
  size_t ws_bytes = 0;
  pseg_plan_workspace_size(P, 1, samples, &ws_bytes);
  // the row, row_len, every row's length, the keys, one factor per
  // calibration step, the workspace
  const size_t off_len = ((size_t)samples * esz + 255) / 256 * 256;
  const size_t off_keys = off_len + 256, off_k = off_keys + 256, off_ws = off_k + 256;
  char *dev = nullptr;
  hipError_t e = hipMalloc((void **)&dev, off_ws + ws_bytes + 256);
  if (e != hipSuccess) {
    pseg_plan_destroy(P);
    return hip_fail(e, "hipMalloc of the impulse row");
  }
  auto done = [&](pseg_status s) {
    (void)hipFree(dev);
    pseg_plan_destroy(P);
    return s;
  };
  const int32_t len = samples;
  const double keys[2] = {0.0, 0.0};
  double k[2] = {1.0, 1.0};
  if (P->n_calib > 0) {
    // the impulse row's keys: live, no sensor, channel or station, no file scales
    pseg_calib_row row;
    memset(&row, 0, sizeof row);
    row.channel_set = -1;
    row.channel = row.rcv_line = row.rcv_point = NAN;
    row.file_descale = row.file_sensitivity = NAN;
    st = pseg_plan_calib_factors(P, &row, 1, 0, k, nullptr, nullptr);
    if (st != PSEG_OK) return done(st);
  }
  e = hipMemcpy(dev, host.data(), host.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dev + off_len, &len, sizeof len, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dev + off_keys, keys, sizeof keys, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dev + off_k, k, sizeof k, hipMemcpyHostToDevice);
  if (e != hipSuccess) return done(hip_fail(e, "copy of the impulse row"));
  st = pseg_plan_apply_calib(P, dev, dev, 1, samples, (const int32_t *)(dev + off_len),
                             (const double *)(dev + off_keys), 0, (const int32_t *)(dev + off_len),
                             (const double *)(dev + off_k), ws_bytes ? dev + off_ws : nullptr,
                             ws_bytes, nullptr);
  if (st != PSEG_OK) return done(st);
  e = hipMemcpy(host.data(), dev, host.size(), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return done(hip_fail(e, "copy of the impulse response"));
  std::vector<double> x(samples);
  for (int32_t j = 0; j < samples; j++)
    x[j] = dtype == PSEG_F32 ? (double)((float *)host.data())[j] : ((double *)host.data())[j];
  return done(pseg_response_moduli(x.data(), samples, hz, m, dt_ms / 1000, gain));
}

}  // extern "C"
