// gfx950 kernels of the processSEG panel hot path.
//
// pipeline_kernel: one workgroup (256 threads, 4 wave64) per row.  The row
// crosses HBM once: it is loaded into registers (sample j on thread j mod 256,
// slot j / 256, so every load and store is coalesced), every step of the plan
// runs on it there and in LDS, and it is stored once.  Steps:
//   demean / detrend / scale / clip / polarity   core/basic.nim:52-110
//   FFT filter (trend, pads, rFFT, gain, irFFT)  filter/apply.nim:194-363
//   AGC (block prefix/suffix sums, gain step)     gain/gain.nim:351-461
//   per-trace normalization (max, rms, exact percentile by radix select)
//                                                 gain/gain.nim:198-219, gain/stats.nim:159-182
// The real FFT of n points is a complex Stockham FFT of m = n/2 points on the
// interleaved samples (fft/plan.nim:501-594); its stages ping-pong between two
// LDS buffers of m complex values.  Split, gain and the inverse's unsplit are
// one fused pass over the bins, and the inverse transform is the forward one
// on swapped (im, re) pairs, so the swap back is an XOR of the sample index.
// Sums (demean, trend, rms, AGC windows) run in float64 as in the reference;
// the AGC window sums are block prefix and suffix sums without subtraction.
//
// corr_kernel: one workgroup (1024 threads) per record: gather to the live
// length, rFFT, times the device-resident pilot multiplier, irFFT, lag spans
// (vibroseis/correlate.nim:241-290).
//
// panel kernels: per-row sums of squares and maxima, a 65536-bin histogram of
// float32 |x| bits >> 15 with integer atomics (bit-exact counts), and a
// one-block kernel for the per-panel factor (gain/stats.nim:253-304,
// gain/gain.nim:221-240).
#include "pseg_internal.hpp"

#include <math.h>

namespace pseg {

namespace {

constexpr int NT = 256;           // threads of the pipeline kernel
constexpr int NW = NT / 64;       // waves of the pipeline kernel

template <typename T> struct V2T;
template <> struct V2T<float> { using type = float2; };
template <> struct V2T<double> { using type = double2; };

template <typename T> using V2 = typename V2T<T>::type;

template <typename T>
__device__ __forceinline__ V2<T> mk(T x, T y) {
  V2<T> r;
  r.x = x;
  r.y = y;
  return r;
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, FastDiv d) {
  return (__umulhi(n, d.mul) + n) >> d.shift;
}

// --------------------------------------------------------------- reductions

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// The sum of v over the block, the same value on every thread, in a fixed
// order (lane 0 of each wave, then the waves in order): a row's result does
// not depend on anything but the row.
template <int NTH>
__device__ double block_sum(double v, double *red) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0;
#pragma unroll
  for (int i = 0; i < NTH / 64; i++) t += red[i];
  return t;
}

template <int NTH>
__device__ void block_sum2(double &a, double &b, double *red) {
  a = wave_sum(a);
  b = wave_sum(b);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6] = a;
    red[NTH / 64 + (threadIdx.x >> 6)] = b;
  }
  __syncthreads();
  double ta = 0, tb = 0;
#pragma unroll
  for (int i = 0; i < NTH / 64; i++) {
    ta += red[i];
    tb += red[NTH / 64 + i];
  }
  a = ta;
  b = tb;
}

template <typename U>
__device__ __forceinline__ U wave_max_u(U v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    U w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}

// ------------------------------------------------------------ bit helpers

template <typename T> struct Bits;
template <> struct Bits<float> {
  using S = int32_t;
  using U = uint32_t;
  static constexpr int width = 32;
  __device__ static S abs_bits(float x) { return (S)(__float_as_uint(x) & 0x7FFFFFFFu); }
  __device__ static double value(S b) { return (double)__uint_as_float((uint32_t)b); }
};
template <> struct Bits<double> {
  using S = long long;
  using U = unsigned long long;
  static constexpr int width = 64;
  __device__ static S abs_bits(double x) {
    return (S)(__double_as_longlong(x) & 0x7FFFFFFFFFFFFFFFll);
  }
  __device__ static double value(S b) { return __longlong_as_double(b); }
};

// fitsIn (gain/gain.nim:112-118): finite and positive, also in T.
template <typename T>
__device__ __forceinline__ bool fits_in(double x) {
  if (!(isfinite(x) && x > 0)) return false;
  double t = (double)(T)x;
  return isfinite(t) && t > 0;
}

// --------------------------------------------------------------- FFT stage

template <typename T>
__device__ __forceinline__ V2<T> cmul(V2<T> a, V2<T> w) {
  return mk<T>(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
}

// Radix-R butterflies of the Stockham DIF stage (fft/plan.nim:225-366),
// in place on v[0 .. R).
template <typename T, int R> struct Bfly;

template <typename T> struct Bfly<T, 2> {
  __device__ static void run(V2<T> *v) {
    V2<T> u = v[0], w = v[1];
    v[0] = mk<T>(u.x + w.x, u.y + w.y);
    v[1] = mk<T>(u.x - w.x, u.y - w.y);
  }
};

template <typename T> struct Bfly<T, 3> {
  __device__ static void run(V2<T> *v) {
    const T c = (T)(-0.5);
    const T s = (T)0.8660254037844386467637231707529362;
    V2<T> a = v[0];
    T t1r = v[1].x + v[2].x, t1i = v[1].y + v[2].y;
    T t2r = v[1].x - v[2].x, t2i = v[1].y - v[2].y;
    T m1r = a.x + c * t1r, m1i = a.y + c * t1i;
    T m2r = s * t2i, m2i = -s * t2r;
    v[0] = mk<T>(a.x + t1r, a.y + t1i);
    v[1] = mk<T>(m1r + m2r, m1i + m2i);
    v[2] = mk<T>(m1r - m2r, m1i - m2i);
  }
};

template <typename T> struct Bfly<T, 4> {
  __device__ static void run(V2<T> *v) {
    T t0r = v[0].x + v[2].x, t0i = v[0].y + v[2].y;
    T t1r = v[0].x - v[2].x, t1i = v[0].y - v[2].y;
    T t2r = v[1].x + v[3].x, t2i = v[1].y + v[3].y;
    T t3r = v[1].y - v[3].y, t3i = v[3].x - v[1].x;  // (x1 - x3)(-i)
    v[0] = mk<T>(t0r + t2r, t0i + t2i);
    v[1] = mk<T>(t1r + t3r, t1i + t3i);
    v[2] = mk<T>(t0r - t2r, t0i - t2i);
    v[3] = mk<T>(t1r - t3r, t1i - t3i);
  }
};

template <typename T> struct Bfly<T, 5> {
  __device__ static void run(V2<T> *v) {
    const T c1 = (T)0.3090169943749474241022934171828191;
    const T c2 = (T)(-0.8090169943749474241022934171828191);
    const T s1 = (T)0.9510565162951535721164393333793821;
    const T s2 = (T)0.5877852522924731291687059546390728;
    V2<T> a = v[0];
    T t1r = v[1].x + v[4].x, t1i = v[1].y + v[4].y;
    T t2r = v[2].x + v[3].x, t2i = v[2].y + v[3].y;
    T t3r = v[1].x - v[4].x, t3i = v[1].y - v[4].y;
    T t4r = v[2].x - v[3].x, t4i = v[2].y - v[3].y;
    T b1r = a.x + c1 * t1r + c2 * t2r, b1i = a.y + c1 * t1i + c2 * t2i;
    T b2r = a.x + c2 * t1r + c1 * t2r, b2i = a.y + c2 * t1i + c1 * t2i;
    T d1r = s1 * t3r + s2 * t4r, d1i = s1 * t3i + s2 * t4i;
    T d2r = s2 * t3r - s1 * t4r, d2i = s2 * t3i - s1 * t4i;
    v[0] = mk<T>(a.x + t1r + t2r, a.y + t1i + t2i);
    v[1] = mk<T>(b1r + d1i, b1i - d1r);
    v[4] = mk<T>(b1r - d1i, b1i + d1r);
    v[2] = mk<T>(b2r + d2i, b2i - d2r);
    v[3] = mk<T>(b2r - d2i, b2i + d2r);
  }
};

// One Stockham stage x -> y: y[q + s(rp + k)] = (sum_j x[q + s(p + j sub)]
// w_r^{jk}) w_cur^{pk}, butterfly b = p s + q.
template <typename T, int R, int NTH>
__device__ __forceinline__ void stage(const V2<T> *__restrict__ x, V2<T> *__restrict__ y,
                                      const V2<T> *__restrict__ tw, int m, int s, int sub,
                                      FastDiv sd) {
  const int nb = m / R;
  const bool twiddle = sub > 1;
  for (int b = threadIdx.x; b < nb; b += NTH) {
    int p = (int)fdiv((uint32_t)b, sd);
    int q = b - p * s;
    V2<T> v[R];
#pragma unroll
    for (int j = 0; j < R; j++) v[j] = x[q + s * (p + j * sub)];
    Bfly<T, R>::run(v);
    if (twiddle) {
#pragma unroll
      for (int k = 1; k < R; k++) v[k] = cmul<T>(v[k], tw[p * (R - 1) + k - 1]);
    }
#pragma unroll
    for (int k = 0; k < R; k++) y[q + s * (R * p + k)] = v[k];
  }
}

// The complex forward FFT of a (m points) with b as the other buffer;
// returns the buffer holding the result.
template <typename T, int NTH>
__device__ V2<T> *complex_forward(const DevFft &f, V2<T> *a, V2<T> *b) {
  const V2<T> *tw = (const V2<T> *)f.tw;
  V2<T> *x = a, *y = b;
  int s = 1;
  for (int i = 0; i < f.nrad; i++) {
    const int r = f.rad[i], sub = f.sub[i];
    const V2<T> *t = tw + f.tw_off[i];
    switch (r) {
      case 4: stage<T, 4, NTH>(x, y, t, f.m, s, sub, f.sdiv[i]); break;
      case 2: stage<T, 2, NTH>(x, y, t, f.m, s, sub, f.sdiv[i]); break;
      case 3: stage<T, 3, NTH>(x, y, t, f.m, s, sub, f.sdiv[i]); break;
      default: stage<T, 5, NTH>(x, y, t, f.m, s, sub, f.sdiv[i]); break;
    }
    __syncthreads();
    s *= r;
    V2<T> *tmp = x;
    x = y;
    y = tmp;
  }
  return x;
}

// splitForward (fft/plan.nim:470-482): bin k of the real transform from
// Z[k] = a and Z[m-k] = b.
template <typename T>
__device__ __forceinline__ V2<T> split_fwd(V2<T> a, V2<T> b, V2<T> w) {
  T er = (T)0.5 * (a.x + b.x);
  T ei = (T)0.5 * (a.y - b.y);
  T orr = (T)0.5 * (a.y + b.y);
  T oi = (T)0.5 * (b.x - a.x);
  return mk<T>(er + w.x * orr - w.y * oi, ei + w.x * oi + w.y * orr);
}

// splitInverse (fft/plan.nim:484-499) with w = (cos, -sin): the swapped
// (Im Z', Re Z') of the inverse's input from X[k] = a and X[m-k] = b.
template <typename T>
__device__ __forceinline__ V2<T> split_inv(V2<T> a, V2<T> b, V2<T> w, T h) {
  T er = a.x + b.x, ei = a.y - b.y, dr = a.x - b.x, di = a.y + b.y;
  T wr = w.x, wi = -w.y;
  T orr = dr * wr - di * wi;
  T oi = dr * wi + di * wr;
  return mk<T>((ei + orr) * h, (er - oi) * h);
}

// Bin multipliers of the fused split / multiply / unsplit pass.
template <typename T> struct GainMul {
  const T *g;
  __device__ V2<T> operator()(V2<T> x, int k) const {
    T w = g[k];
    return mk<T>(x.x * w, x.y * w);
  }
};
template <typename T> struct CplxMul {
  const V2<T> *mul;
  __device__ V2<T> operator()(V2<T> x, int k) const {
    V2<T> w = mul[k];
    return mk<T>(x.x * w.x - x.y * w.y, x.x * w.y + x.y * w.x);
  }
};

// Z (the forward complex FFT of the packed samples) -> W, the swapped input
// of the inverse: split into bins, multiply, unsplit with the 1/n scale.
template <typename T, int NTH, typename Mul>
__device__ void combine(const DevFft &f, const V2<T> *__restrict__ Z, V2<T> *__restrict__ W,
                        Mul mul) {
  const int m = f.m;
  const V2<T> *sp = (const V2<T> *)f.split;
  const T scale = (T)1 / (T)f.n;
  for (int k = threadIdx.x; k <= m / 2; k += NTH) {
    if (k == 0) {
      V2<T> z = Z[0];
      V2<T> x0 = mul(mk<T>(z.x + z.y, (T)0), 0);
      V2<T> xm = mul(mk<T>(z.x - z.y, (T)0), m);
      W[0] = mk<T>((x0.x - xm.x) * scale, (x0.x + xm.x) * scale);
    } else {
      const int kk = m - k;
      V2<T> a = Z[k], b = Z[kk];
      V2<T> xk = mul(split_fwd<T>(a, b, sp[k]), k);
      if (kk != k) {
        V2<T> xkk = mul(split_fwd<T>(b, a, sp[kk]), kk);
        W[k] = split_inv<T>(xk, xkk, sp[k], scale);
        W[kk] = split_inv<T>(xkk, xk, sp[kk], scale);
      } else {
        W[k] = split_inv<T>(xk, xk, sp[k], scale);
      }
    }
  }
}

// ------------------------------------------------------------ row helpers

template <typename T, int PER>
__device__ __forceinline__ double row_sum(const T (&v)[PER], int n) {
  double s = 0;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    int j = threadIdx.x + NT * k;
    if (j < n) s += (double)v[k];
  }
  return s;
}

// Least-squares line (filter/edges.nim:16-30, core/basic.nim:61-81 share it).
template <typename T, int PER>
__device__ void trend_of(const T (&v)[PER], int n, int trend, double *red, double &a,
                         double &b) {
  if (trend == 0) {
    a = 0;
    b = 0;
    return;
  }
  double sx = 0, sxi = 0;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    int j = threadIdx.x + NT * k;
    if (j < n) {
      sx += (double)v[k];
      sxi += (double)j * (double)v[k];
    }
  }
  if (trend == 1) {
    a = block_sum<NT>(sx, red) / (double)n;
    b = 0;
    return;
  }
  block_sum2<NT>(sx, sxi, red);
  double nf = (double)n, mean = sx / nf;
  if (n == 1) {
    a = mean;
    b = 0;
    return;
  }
  double c = (nf - 1) / 2;
  double slope = (sxi - c * sx) / (nf * (nf * nf - 1) / 12);
  a = mean - slope * c;
  b = slope;
}

// One FFT filter step (filter/apply.nim:340-363) on the row in registers.
template <typename T, int PER>
__device__ void filter_step(const DevFilter &F, T (&v)[PER], int n, char *lds, double *red) {
  const DevFft &f = F.fft;
  T *buf = (T *)lds;                          // nfft samples = m complex
  V2<T> *A = (V2<T> *)lds;
  V2<T> *B = (V2<T> *)(lds + sizeof(T) * (size_t)f.n);
  double a, b;
  trend_of<T, PER>(v, n, F.trend, red, a, b);
  const int p = F.pad, nf = f.n;
  // pack (packInto, filter/apply.nim:194-216): detrended data ...
#pragma unroll
  for (int k = 0; k < PER; k++) {
    int j = threadIdx.x + NT * k;
    if (j < n) buf[p + j] = (T)((double)v[k] - (a + b * (double)j));
  }
  if (!F.reflect) {
    for (int j = threadIdx.x; j < p; j += NT) buf[j] = (T)0;
    for (int j = p + n + threadIdx.x; j < nf; j += NT) buf[j] = (T)0;
  } else {
    for (int j = 2 * p + n + threadIdx.x; j < nf; j += NT) buf[j] = (T)0;
  }
  __syncthreads();
  if (F.reflect) {
    // ... and odd reflections 2 x_edge - x_k tapered by fall[k - 1]
    const T *fall = (const T *)F.fall;
    const T d0 = buf[p], dn = buf[p + n - 1];
    for (int k = 1 + threadIdx.x; k <= p; k += NT) {
      int mk_ = k < n - 1 ? k : n - 1;
      T fk = fall[k - 1];
      T rl = buf[p + mk_];
      T rr = buf[p + n - 1 - mk_];
      buf[p - k] = (d0 + d0 - rl) * fk;
      buf[p + n - 1 + k] = (dn + dn - rr) * fk;
    }
    __syncthreads();
  }
  V2<T> *Z = complex_forward<T, NT>(f, A, B);
  V2<T> *W = Z == A ? B : A;
  combine<T, NT>(f, Z, W, GainMul<T>{(const T *)F.gain});
  __syncthreads();
  V2<T> *R = complex_forward<T, NT>(f, W, Z);
  const T *out = (const T *)R;
  // unpack (unpackFrom, filter/apply.nim:218-231): sample i of the inverse is
  // element i ^ 1 of the swapped result.
  double gb = b * F.dc_gain;
  double ga = F.zero_dc ? -gb * (double)(n - 1) / 2 : a * F.dc_gain;
  const bool plain = ga == 0 && gb == 0;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    int j = threadIdx.x + NT * k;
    if (j < n) {
      T y = out[(p + j) ^ 1];
      v[k] = plain ? y : (T)((double)y + (ga + gb * (double)j));
    }
  }
  __syncthreads();
}

// Inclusive scan over the 64 lanes of a wave, from lane 0 up (dir = 0) or
// from lane 63 down (dir = 1).
__device__ __forceinline__ double wave_scan(double v, int dir) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    double o = dir == 0 ? __shfl_up(v, d, 64) : __shfl_down(v, d, 64);
    bool take = dir == 0 ? lane >= d : lane + d < 64;
    if (take) v += o;
  }
  return v;
}

// AGC (gain/gain.nim:404-461) on the row in registers.
template <typename T, int PER>
__device__ void agc_step(const DevAgc &G, T (&v)[PER], int n, char *lds, double *red) {
  const int mode = G.mode;
  double sq = 0;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    int j = threadIdx.x + NT * k;
    if (j < n) sq += (double)v[k] * (double)v[k];
  }
  double rms = sqrt(block_sum<NT>(sq, red) / (double)n);
  if (rms == 0) return;  // all zero: nothing to scale
  if (!fits_in<T>(G.target / rms / G.epsilon)) return;
  const int h = G.half, w = 2 * h + 1;
  T *s = (T *)lds;          // x, then the suffix sums in place
  T *pp = s + n;            // prefix sums, n + 1 values
#pragma unroll
  for (int k = 0; k < PER; k++) {
    int j = threadIdx.x + NT * k;
    if (j < n) s[j] = v[k];
  }
  __syncthreads();
  // blockSums (gain/gain.nim:351-390): per block of w samples, each wave one
  // block at a time, each lane a run of c samples; float64 sums of
  // non-negative terms only.
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nblk = (n + w - 1) / w;
  for (int blk = wave; blk < nblk; blk += NW) {
    const int bs = blk * w, be = min(bs + w, n), len = be - bs;
    const int c = (len + 63) / 64;
    const int c0 = min(bs + lane * c, be), c1 = min(c0 + c, be);
    double tot = 0;
    for (int i = c0; i < c1; i++) {
      double x = (double)s[i];
      tot += mode == 0 ? x * x : fabs(x);
    }
    double incl = wave_scan(tot, 0);
    double excl = __shfl_up(incl, 1, 64);
    if (lane == 0) excl = 0;
    double rincl = wave_scan(tot, 1);
    double rexcl = __shfl_down(rincl, 1, 64);
    if (lane == 63) rexcl = 0;
    if (lane == 0) pp[bs] = (T)0;
    double run = excl;
    for (int i = c0; i < c1; i++) {
      double x = (double)s[i];
      run += mode == 0 ? x * x : fabs(x);
      // p[i + 1]; a block's end is the next block's start (0), except at n
      if (i + 1 < be || (be == n && n % w != 0)) pp[i + 1] = (T)run;
    }
    run = rexcl;
    for (int i = c1 - 1; i >= c0; i--) {
      double x = (double)s[i];
      run += mode == 0 ? x * x : fabs(x);
      s[i] = (T)run;
    }
  }
  if (threadIdx.x == 0 && n % w == 0) pp[n] = (T)0;
  __syncthreads();
  // The gain step in T on means scaled by the trace level, five ranges of i.
  const double unit = mode == 0 ? 1 / (rms * rms) : 1 / rms;
  const T kk = (T)(G.target / rms);
  const T eps = (T)G.epsilon;
  const int first = min(h, n);
  const int last = max(first, n - h);
  const int headCut = min(max(n - h, 0), first);
  const int lastBlock = (n - 1) / w * w;
  const int tailCut = min(max(lastBlock + h, last), n);
  const double pn = (double)pp[n];
  const double *inv = G.inv, *rev = G.rev;
  const double unitW = unit / (double)w;
  const double headWhole = headCut < first ? pn * unit * inv[n] : 0.0;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    int i = threadIdx.x + NT * k;
    if (i < n) {
      double mean;
      if (i < headCut) mean = (double)pp[i + h + 1] * (unit * inv[i + h + 1]);
      else if (i < first) mean = headWhole;
      else if (i < last) mean = ((double)s[i - h] + (double)pp[i + h + 1]) * unitW;
      else if (i < tailCut) mean = ((double)s[i - h] + pn) * (unit * rev[i + h + 1 - n]);
      else mean = (double)s[i - h] * (unit * rev[i + h + 1 - n]);
      T lvl = mode == 0 ? (T)sqrt((T)mean) : (T)mean;
      T gi = kk / (lvl + eps);
      v[k] = (T)(v[k] * gi);
    }
  }
  __syncthreads();
}

// The k-th smallest of the row's |x| bit patterns (exact), by radix select
// over 8-bit digits with an LDS histogram.
template <typename T, int PER>
__device__ typename Bits<T>::S select_bits(const T (&v)[PER], int n, int64_t rank,
                                           uint32_t *hist, int64_t *shared_i) {
  using S = typename Bits<T>::S;
  using U = typename Bits<T>::U;
  U prefix = 0, mask = 0;
  int64_t r = rank;
  for (int shift = Bits<T>::width - 8; shift >= 0; shift -= 8) {
    for (int i = threadIdx.x; i < 256; i += NT) hist[i] = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; k++) {
      int j = threadIdx.x + NT * k;
      if (j < n) {
        U key = (U)Bits<T>::abs_bits(v[k]);
        if ((key & mask) == prefix) atomicAdd(&hist[(key >> shift) & 255u], 1u);
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int64_t below = 0;
      int d = 0;
      for (; d < 256; d++) {
        int64_t c = hist[d];
        if (below + c > r) break;
        below += c;
      }
      shared_i[0] = d;
      shared_i[1] = r - below;
    }
    __syncthreads();
    prefix |= (U)shared_i[0] << shift;
    mask |= (U)255 << shift;
    r = shared_i[1];
    __syncthreads();
  }
  return (S)prefix;
}

// Per-trace normalization (gain/gain.nim:198-219).
template <typename T, int PER>
__device__ void norm_trace_step(const DevStep &st, T (&v)[PER], int n, char *lds,
                                double *red) {
  using S = typename Bits<T>::S;
  __shared__ S smax[NW];
  __shared__ int64_t sh_i[2];
  // max |x| as the integer maximum of the sign-cleared bits (a NaN sorts
  // above Inf: gain/stats.nim:97-117)
  S mb = 0;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    int j = threadIdx.x + NT * k;
    if (j < n) {
      S b = Bits<T>::abs_bits(v[k]);
      mb = b > mb ? b : mb;
    }
  }
  mb = wave_max_u<S>(mb);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) smax[threadIdx.x >> 6] = mb;
  __syncthreads();
  mb = smax[0];
  for (int i = 1; i < NW; i++) mb = smax[i] > mb ? smax[i] : mb;
  const double mx = Bits<T>::value(mb);
  double level;
  if (st.stat == 0) {
    level = mx;
  } else if (st.stat == 1) {
    double sq = 0;
#pragma unroll
    for (int k = 0; k < PER; k++) {
      int j = threadIdx.x + NT * k;
      if (j < n) sq += (double)v[k] * (double)v[k];
    }
    level = sqrt(block_sum<NT>(sq, red) / (double)n);
  } else {
    // percentileAbs (gain/stats.nim:159-182): linear between order
    // statistics k and k + 1 of |x|, h = (n - 1) p / 100.
    if (isnan(mx)) return;
    double hh = (double)(n - 1) * st.percentile / 100;
    int64_t k = (int64_t)floor(hh);
    if (k > n - 1) k = n - 1;
    uint32_t *hist = (uint32_t *)lds;
    S lo = select_bits<T, PER>(v, n, k, hist, sh_i);
    double lower = Bits<T>::value(lo);
    if (k + 1 >= n) {
      level = lower;
    } else {
      // the next order statistic: lower again if it is tied, else the
      // smallest value above it
      double cnt = 0;
      S up = (S)0x7FFFFFFFFFFFFFFFll;
      if (sizeof(S) == 4) up = (S)0x7FFFFFFF;
#pragma unroll
      for (int q = 0; q < PER; q++) {
        int j = threadIdx.x + NT * q;
        if (j < n) {
          S b = Bits<T>::abs_bits(v[q]);
          if (b <= lo) cnt += 1;
          else up = b < up ? b : up;
        }
      }
      cnt = block_sum<NT>(cnt, red);
      // block minimum of up
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        S w2 = __shfl_xor(up, o, 64);
        up = w2 < up ? w2 : up;
      }
      __syncthreads();
      if ((threadIdx.x & 63) == 0) smax[threadIdx.x >> 6] = up;
      __syncthreads();
      up = smax[0];
      for (int i = 1; i < NW; i++) up = smax[i] < up ? smax[i] : up;
      double upper = cnt >= (double)(k + 2) ? lower : Bits<T>::value(up);
      double f = hh - (double)k;
      level = f == 0 ? lower : lower + f * (upper - lower);
    }
  }
  // scaleToTarget (gain/gain.nim:208-210)
  if (isfinite(level) && level > 0 && fits_in<T>(st.target / level)) {
    const T f = (T)(st.target / level);
#pragma unroll
    for (int k = 0; k < PER; k++) {
      int j = threadIdx.x + NT * k;
      if (j < n) v[k] = v[k] * f;
    }
  }
  __syncthreads();
}

}  // namespace

// ---------------------------------------------------------------- kernels

template <typename T, int PER>
__global__ __launch_bounds__(NT) void pipeline_kernel(const DevPlan *__restrict__ plan,
                                                      int32_t first, int32_t last,
                                                      int32_t cols, int32_t has_factor,
                                                      const double *__restrict__ factor,
                                                      const T *__restrict__ src,
                                                      T *__restrict__ dst,
                                                      const int32_t *__restrict__ row_len) {
  extern __shared__ __align__(16) char lds[];
  __shared__ double red[2 * NW];
  const int64_t row = blockIdx.x;
  const int n = row_len[row];
  const T *s = src + row * (int64_t)cols;
  T *d = dst + row * (int64_t)cols;
  T v[PER];
#pragma unroll
  for (int k = 0; k < PER; k++) {
    int j = threadIdx.x + NT * k;
    v[k] = j < cols ? s[j] : (T)0;
  }
  if (n > 0) {
    if (has_factor) {
      // a per-panel factor at the start of the part (pipeline.nim:717-718)
      const double f = *factor;
      if (!isnan(f)) {
        const T kf = (T)f;
#pragma unroll
        for (int k = 0; k < PER; k++) {
          int j = threadIdx.x + NT * k;
          if (j < n) v[k] = v[k] * kf;
        }
      }
    }
    for (int i = first; i < last; i++) {
      const DevStep &st = plan->steps[i];
      switch (st.kind) {
        case kDemean: {
          // demean (core/basic.nim:52-59)
          const T m = (T)(block_sum<NT>(row_sum<T, PER>(v, n), red) / (double)n);
#pragma unroll
          for (int k = 0; k < PER; k++) {
            int j = threadIdx.x + NT * k;
            if (j < n) v[k] = v[k] - m;
          }
          break;
        }
        case kDetrend: {
          // detrend (core/basic.nim:61-81)
          double a, b;
          trend_of<T, PER>(v, n, 2, red, a, b);
          const T ta = (T)a, tb = (T)b;
#pragma unroll
          for (int k = 0; k < PER; k++) {
            int j = threadIdx.x + NT * k;
            if (j < n) v[k] = v[k] - (ta + tb * (T)j);
          }
          break;
        }
        case kScale: {
          const T kf = (T)st.factor;
#pragma unroll
          for (int k = 0; k < PER; k++) {
            int j = threadIdx.x + NT * k;
            if (j < n) v[k] = v[k] * kf;
          }
          break;
        }
        case kClip: {
          const T lo = (T)st.lo, hi = (T)st.hi;
#pragma unroll
          for (int k = 0; k < PER; k++) {
            int j = threadIdx.x + NT * k;
            if (j < n) {
              T x = v[k];
              v[k] = x > hi ? hi : (x < lo ? lo : x);
            }
          }
          break;
        }
        case kPolarity: {
#pragma unroll
          for (int k = 0; k < PER; k++) {
            int j = threadIdx.x + NT * k;
            if (j < n) v[k] = -v[k];
          }
          break;
        }
        case kFilter: filter_step<T, PER>(st.filter, v, n, lds, red); break;
        case kAgc: agc_step<T, PER>(st.agc, v, n, lds, red); break;
        case kNormTrace: norm_trace_step<T, PER>(st, v, n, lds, red); break;
        default: break;
      }
    }
  }
#pragma unroll
  for (int k = 0; k < PER; k++) {
    int j = threadIdx.x + NT * k;
    if (j < cols) d[j] = v[k];
  }
}

// Clears n uint32 (a kernel, not hipMemsetAsync, so that a captured graph
// holds no memset node).
__global__ void zero_u32_kernel(uint32_t *__restrict__ p, int32_t n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

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
  const int n = row_len[row];
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
  const int n = row_len[row];
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

// Vibroseis correlation of one record per workgroup.
template <typename T, int NTC>
__global__ __launch_bounds__(NTC) void corr_kernel(const DevCorr *__restrict__ plan,
                                                   const T *__restrict__ src, int32_t src_cols,
                                                   const int32_t *__restrict__ row_len,
                                                   T *__restrict__ dst) {
  extern __shared__ __align__(16) char lds[];
  const DevCorr &c = *plan;
  const int64_t row = blockIdx.x;
  const int live = row_len[row];
  T *d = dst + row * (int64_t)c.out;
  if (live < 0) {
    for (int j = threadIdx.x; j < c.out; j += NTC) d[j] = (T)0;
    return;
  }
  const DevFft &f = c.fft;
  const int n = f.n;
  T *buf = (T *)lds;
  V2<T> *A = (V2<T> *)lds;
  V2<T> *B = (V2<T> *)(lds + sizeof(T) * (size_t)n);
  const T *s = src + row * (int64_t)src_cols;
  for (int j = threadIdx.x; j < n; j += NTC) buf[j] = j < live ? s[j] : (T)0;
  __syncthreads();
  V2<T> *Z = complex_forward<T, NTC>(f, A, B);
  V2<T> *W = Z == A ? B : A;
  combine<T, NTC>(f, Z, W, CplxMul<T>{(const V2<T> *)c.mul});
  __syncthreads();
  const T *out = (const T *)complex_forward<T, NTC>(f, W, Z);
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

// -------------------------------------------------------------- launchers

namespace {

template <typename T, int PER>
hipError_t launch_one(const DevPlan *dplan, const PipeLaunch &a, const T *src, T *dst,
                      int32_t rows, const int32_t *row_len, hipStream_t stream) {
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute((const void *)pipeline_kernel<T, PER>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       160 * 1024 - 2048);
    if (e != hipSuccess) return e;
    attr = true;
  }
  hipLaunchKernelGGL((pipeline_kernel<T, PER>), dim3(rows), dim3(NT), a.lds_bytes, stream,
                     dplan, a.first, a.last, a.cols, a.has_factor, a.factor, src, dst, row_len);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_typed(const DevPlan *dplan, const PipeLaunch &a, const void *src, void *dst,
                        int32_t rows, const int32_t *row_len, hipStream_t stream) {
  const int per = (a.cols + NT - 1) / NT;
  const T *s = (const T *)src;
  T *d = (T *)dst;
  if (per <= 2) return launch_one<T, 2>(dplan, a, s, d, rows, row_len, stream);
  if (per <= 4) return launch_one<T, 4>(dplan, a, s, d, rows, row_len, stream);
  if (per <= 8) return launch_one<T, 8>(dplan, a, s, d, rows, row_len, stream);
  if (per <= 12) return launch_one<T, 12>(dplan, a, s, d, rows, row_len, stream);
  if (per <= 16) return launch_one<T, 16>(dplan, a, s, d, rows, row_len, stream);
  if (per <= 24) return launch_one<T, 24>(dplan, a, s, d, rows, row_len, stream);
  if (per <= 32) return launch_one<T, 32>(dplan, a, s, d, rows, row_len, stream);
  if (per <= 48) return launch_one<T, 48>(dplan, a, s, d, rows, row_len, stream);
  return hipErrorInvalidValue;
}

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

template <typename T>
hipError_t corr_typed(const DevCorr *dplan, size_t lds, const T *src, int32_t rows,
                      int32_t src_cols, const int32_t *row_len, T *dst, hipStream_t stream) {
  static bool attr = false;
  if (!attr) {
    hipError_t e = hipFuncSetAttribute((const void *)corr_kernel<T, 1024>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize,
                                       160 * 1024 - 2048);
    if (e != hipSuccess) return e;
    attr = true;
  }
  hipLaunchKernelGGL((corr_kernel<T, 1024>), dim3(rows), dim3(1024), lds, stream, dplan, src,
                     src_cols, row_len, dst);
  return hipGetLastError();
}

}  // namespace

int pipeline_max_cols() { return NT * 48; }

hipError_t launch_pipeline(int dtype, const DevPlan *dev_plan, const PipeLaunch &a,
                           const void *src, void *dst, int32_t rows, const int32_t *row_len,
                           hipStream_t stream) {
  if (rows <= 0) return hipSuccess;
  if (dtype == 0) return launch_typed<float>(dev_plan, a, src, dst, rows, row_len, stream);
  return launch_typed<double>(dev_plan, a, src, dst, rows, row_len, stream);
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

hipError_t launch_correlate(int dtype, const DevCorr *dev_plan, size_t lds_bytes,
                            const void *src, int32_t rows, int32_t src_cols,
                            const int32_t *row_len, void *dst, hipStream_t stream) {
  if (rows <= 0) return hipSuccess;
  if (dtype == 0)
    return corr_typed<float>(dev_plan, lds_bytes, (const float *)src, rows, src_cols, row_len,
                             (float *)dst, stream);
  return corr_typed<double>(dev_plan, lds_bytes, (const double *)src, rows, src_cols, row_len,
                            (double *)dst, stream);
}

}  // namespace pseg
