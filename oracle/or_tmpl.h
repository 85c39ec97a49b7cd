/*
 * Type-generic body of the oracle (TEST INFRASTRUCTURE ONLY; see oracle.h).
 * oracle.c includes this file twice: with T = double, SFX(x) = x##_f64 and
 * with T = float, SFX(x) = x##_f32.  Tables and sums are float64 in both, as
 * in the reference; butterflies and samples are T.
 */

/* ------------------------------------------------------------- FFT plan -- */

/* FftPlan (fft/plan.nim:92-99, 136-160). */
typedef struct SFX(fftplan) {
  int64_t n, m;
  int32_t nrad;
  int32_t rad[64];
  T *twr, *twi;     /* stage twiddles, stage after stage */
  T *spr, *spi;     /* exp(-2 pi i k / n), k = 0 .. m (even n) */
} SFX(fftplan);

static void SFX(fft_free)(SFX(fftplan) *p) {
  free(p->twr); free(p->twi); free(p->spr); free(p->spi);
  memset(p, 0, sizeof(*p));
}

/* initFftPlan (fft/plan.nim:136-160): twiddles in float64, stored as T. */
static int SFX(fft_init)(SFX(fftplan) *p, int64_t n) {
  memset(p, 0, sizeof(*p));
  if (!or_is_fast_len(n)) return -1;
  p->n = n;
  p->m = (n % 2 == 0) ? n / 2 : n;
  p->nrad = or_factorize(p->m, p->rad);
  int64_t cur = p->m, cnt = 0;
  for (int i = 0; i < p->nrad; i++) {
    int64_t sub = cur / p->rad[i];
    cnt += sub * (p->rad[i] - 1);
    cur = sub;
  }
  p->twr = (T *)malloc(sizeof(T) * (size_t)(cnt + 1));
  p->twi = (T *)malloc(sizeof(T) * (size_t)(cnt + 1));
  cur = p->m;
  int64_t at = 0;
  for (int i = 0; i < p->nrad; i++) {
    int r = p->rad[i];
    int64_t sub = cur / r;
    for (int64_t q = 0; q < sub; q++)
      for (int k = 1; k < r; k++) {
        double a = -2 * OR_PI * (double)((q * k) % cur) / (double)cur;
        p->twr[at] = (T)cos(a);
        p->twi[at] = (T)sin(a);
        at++;
      }
    cur = sub;
  }
  if (n % 2 == 0) {
    p->spr = (T *)malloc(sizeof(T) * (size_t)(p->m + 1));
    p->spi = (T *)malloc(sizeof(T) * (size_t)(p->m + 1));
    for (int64_t k = 0; k <= p->m; k++) {
      double a = -2 * OR_PI * (double)k / (double)n;
      p->spr[k] = (T)cos(a);
      p->spi[k] = (T)sin(a);
    }
  }
  return 0;
}

/* Stockham DIF stage (fft/plan.nim:185-417), one lane:
 *   y[q + s(rp + k)] = (sum_j x[q + s(p + j sub)] w_r^{jk}) w_cur^{pk}. */
#define SFX_MULTW(yr, yi, wr, wi) do { T t_ = (yr) * (wr) - (yi) * (wi); \
    (yi) = (yr) * (wi) + (yi) * (wr); (yr) = t_; } while (0)

static void SFX(pass2)(const T *xr, const T *xi, T *yr, T *yi, const T *tr, const T *ti,
                       int64_t sub, int64_t w) {
  for (int64_t p = 0; p < sub; p++) {
    int64_t a = p * w, b = 2 * p * w, o = sub * w;
    T wr = tr[p], wi = ti[p];
    for (int64_t i = 0; i < w; i++) {
      T ur = xr[a + i], ui = xi[a + i], vr = xr[a + o + i], vi = xi[a + o + i];
      yr[b + i] = ur + vr;
      yi[b + i] = ui + vi;
      T dr = ur - vr, di = ui - vi;
      SFX_MULTW(dr, di, wr, wi);
      yr[b + w + i] = dr;
      yi[b + w + i] = di;
    }
  }
}

static void SFX(pass3)(const T *xr, const T *xi, T *yr, T *yi, const T *tr, const T *ti,
                       int64_t sub, int64_t w) {
  const T c = (T)(-0.5);
  const T s = (T)0.8660254037844386467637231707529362;
  for (int64_t p = 0; p < sub; p++) {
    int64_t a = p * w, b = 3 * p * w, o = sub * w, t = 2 * p;
    T w1r = tr[t], w1i = ti[t], w2r = tr[t + 1], w2i = ti[t + 1];
    for (int64_t i = 0; i < w; i++) {
      T ar = xr[a + i], ai = xi[a + i];
      T t1r = xr[a + o + i] + xr[a + 2 * o + i];
      T t1i = xi[a + o + i] + xi[a + 2 * o + i];
      T t2r = xr[a + o + i] - xr[a + 2 * o + i];
      T t2i = xi[a + o + i] - xi[a + 2 * o + i];
      T m1r = ar + c * t1r, m1i = ai + c * t1i;
      T m2r = s * t2i, m2i = -s * t2r;
      yr[b + i] = ar + t1r;
      yi[b + i] = ai + t1i;
      T u1r = m1r + m2r, u1i = m1i + m2i, u2r = m1r - m2r, u2i = m1i - m2i;
      SFX_MULTW(u1r, u1i, w1r, w1i);
      SFX_MULTW(u2r, u2i, w2r, w2i);
      yr[b + w + i] = u1r; yi[b + w + i] = u1i;
      yr[b + 2 * w + i] = u2r; yi[b + 2 * w + i] = u2i;
    }
  }
}

static void SFX(pass4)(const T *xr, const T *xi, T *yr, T *yi, const T *tr, const T *ti,
                       int64_t sub, int64_t w) {
  for (int64_t p = 0; p < sub; p++) {
    int64_t a = p * w, b = 4 * p * w, o = sub * w, t = 3 * p;
    T w1r = tr[t], w1i = ti[t], w2r = tr[t + 1], w2i = ti[t + 1], w3r = tr[t + 2],
      w3i = ti[t + 2];
    for (int64_t i = 0; i < w; i++) {
      T x0r = xr[a + i], x0i = xi[a + i], x1r = xr[a + o + i], x1i = xi[a + o + i];
      T x2r = xr[a + 2 * o + i], x2i = xi[a + 2 * o + i];
      T x3r = xr[a + 3 * o + i], x3i = xi[a + 3 * o + i];
      T t0r = x0r + x2r, t0i = x0i + x2i, t1r = x0r - x2r, t1i = x0i - x2i;
      T t2r = x1r + x3r, t2i = x1i + x3i;
      T t3r = x1i - x3i, t3i = x3r - x1r; /* (x1 - x3)(-i) */
      yr[b + i] = t0r + t2r;
      yi[b + i] = t0i + t2i;
      T u1r = t1r + t3r, u1i = t1i + t3i, u2r = t0r - t2r, u2i = t0i - t2i;
      T u3r = t1r - t3r, u3i = t1i - t3i;
      SFX_MULTW(u1r, u1i, w1r, w1i);
      SFX_MULTW(u2r, u2i, w2r, w2i);
      SFX_MULTW(u3r, u3i, w3r, w3i);
      yr[b + w + i] = u1r; yi[b + w + i] = u1i;
      yr[b + 2 * w + i] = u2r; yi[b + 2 * w + i] = u2i;
      yr[b + 3 * w + i] = u3r; yi[b + 3 * w + i] = u3i;
    }
  }
}

static void SFX(pass5)(const T *xr, const T *xi, T *yr, T *yi, const T *tr, const T *ti,
                       int64_t sub, int64_t w) {
  const T c1 = (T)0.3090169943749474241022934171828191;
  const T c2 = (T)(-0.8090169943749474241022934171828191);
  const T s1 = (T)0.9510565162951535721164393333793821;
  const T s2 = (T)0.5877852522924731291687059546390728;
  for (int64_t p = 0; p < sub; p++) {
    int64_t a = p * w, b = 5 * p * w, o = sub * w, t = 4 * p;
    T w1r = tr[t], w1i = ti[t], w2r = tr[t + 1], w2i = ti[t + 1];
    T w3r = tr[t + 2], w3i = ti[t + 2], w4r = tr[t + 3], w4i = ti[t + 3];
    for (int64_t i = 0; i < w; i++) {
      T ar = xr[a + i], ai = xi[a + i];
      T x1r = xr[a + o + i], x1i = xi[a + o + i], x2r = xr[a + 2 * o + i],
        x2i = xi[a + 2 * o + i];
      T x3r = xr[a + 3 * o + i], x3i = xi[a + 3 * o + i], x4r = xr[a + 4 * o + i],
        x4i = xi[a + 4 * o + i];
      T t1r = x1r + x4r, t1i = x1i + x4i, t2r = x2r + x3r, t2i = x2i + x3i;
      T t3r = x1r - x4r, t3i = x1i - x4i, t4r = x2r - x3r, t4i = x2i - x3i;
      T b1r = ar + c1 * t1r + c2 * t2r, b1i = ai + c1 * t1i + c2 * t2i;
      T b2r = ar + c2 * t1r + c1 * t2r, b2i = ai + c2 * t1i + c1 * t2i;
      T d1r = s1 * t3r + s2 * t4r, d1i = s1 * t3i + s2 * t4i;
      T d2r = s2 * t3r - s1 * t4r, d2i = s2 * t3i - s1 * t4i;
      yr[b + i] = ar + t1r + t2r;
      yi[b + i] = ai + t1i + t2i;
      T u1r = b1r + d1i, u1i = b1i - d1r, u4r = b1r - d1i, u4i = b1i + d1r;
      T u2r = b2r + d2i, u2i = b2i - d2r, u3r = b2r - d2i, u3i = b2i + d2r;
      SFX_MULTW(u1r, u1i, w1r, w1i);
      SFX_MULTW(u2r, u2i, w2r, w2i);
      SFX_MULTW(u3r, u3i, w3r, w3i);
      SFX_MULTW(u4r, u4i, w4r, w4i);
      yr[b + w + i] = u1r; yi[b + w + i] = u1i;
      yr[b + 2 * w + i] = u2r; yi[b + 2 * w + i] = u2i;
      yr[b + 3 * w + i] = u3r; yi[b + 3 * w + i] = u3i;
      yr[b + 4 * w + i] = u4r; yi[b + 4 * w + i] = u4i;
    }
  }
}

/* complexForward (fft/plan.nim:421-450).  Returns 1 when the result is in b. */
static int SFX(complex_forward)(const SFX(fftplan) *p, T *ar, T *ai, T *br, T *bi) {
  T *xr = ar, *xi = ai, *yr = br, *yi = bi;
  int64_t cur = p->m, s = 1, at = 0;
  for (int i = 0; i < p->nrad; i++) {
    int r = p->rad[i];
    int64_t sub = cur / r;
    int64_t w = s;
    const T *tr = p->twr + at, *ti = p->twi + at;
    switch (r) {
      case 4: SFX(pass4)(xr, xi, yr, yi, tr, ti, sub, w); break;
      case 2: SFX(pass2)(xr, xi, yr, yi, tr, ti, sub, w); break;
      case 3: SFX(pass3)(xr, xi, yr, yi, tr, ti, sub, w); break;
      default: SFX(pass5)(xr, xi, yr, yi, tr, ti, sub, w); break;
    }
    at += sub * (r - 1);
    cur = sub;
    s *= r;
    T *t = xr; xr = yr; yr = t;
    t = xi; xi = yi; yi = t;
  }
  return xr == br;
}

/* forwardImpl (fft/plan.nim:501-542); work holds 4m values. */
static void SFX(fft_forward)(const SFX(fftplan) *p, const T *x, T *re, T *im, T *work) {
  int64_t m = p->m;
  T *ar = work, *ai = work + m, *br = work + 2 * m, *bi = work + 3 * m;
  if (p->n % 2 == 1) {
    for (int64_t i = 0; i < m; i++) { ar[i] = x[i]; ai[i] = (T)0; }
  } else {
    for (int64_t k = 0; k < m; k++) { ar[k] = x[2 * k]; ai[k] = x[2 * k + 1]; }
  }
  int inB = SFX(complex_forward)(p, ar, ai, br, bi);
  T *zr = inB ? br : ar, *zi = inB ? bi : ai;
  int64_t bins = p->n / 2 + 1;
  if (p->n % 2 == 1) {
    for (int64_t k = 0; k < bins; k++) { re[k] = zr[k]; im[k] = zi[k]; }
    return;
  }
  T r0 = zr[0], i0 = zi[0];
  re[0] = r0 + i0; im[0] = (T)0;
  re[m] = r0 - i0; im[m] = (T)0;
  for (int64_t k = 1; k < m; k++) {
    /* splitForward (fft/plan.nim:470-482) */
    T a_r = zr[k], a_i = zi[k], b_r = zr[m - k], b_i = zi[m - k];
    T wr = p->spr[k], wi = p->spi[k];
    T er = (T)0.5 * (a_r + b_r);
    T ei = (T)0.5 * (a_i - b_i);
    T orr = (T)0.5 * (a_i + b_i);
    T oi = (T)0.5 * (b_r - a_r);
    re[k] = er + wr * orr - wi * oi;
    im[k] = ei + wr * oi + wi * orr;
  }
}

/* inverseImpl (fft/plan.nim:544-594): swap(fft(swap(Z))), scaled 1/n. */
static void SFX(fft_inverse)(const SFX(fftplan) *p, const T *re, const T *im, T *x, T *work) {
  int64_t m = p->m;
  T *ar = work, *ai = work + m, *br = work + 2 * m, *bi = work + 3 * m;
  T scale = (T)1 / (T)p->n;
  if (p->n % 2 == 1) {
    int64_t bins = p->n / 2 + 1;
    ai[0] = re[0] * scale;
    ar[0] = (T)0;
    for (int64_t k = 1; k < bins; k++) {
      T vr = re[k] * scale, vi = im[k] * scale;
      ai[k] = vr; ar[k] = vi;
      ai[m - k] = vr; ar[m - k] = -vi;
    }
  } else {
    T x0 = re[0], xm = re[m];
    ai[0] = (x0 + xm) * scale;
    ar[0] = (x0 - xm) * scale;
    for (int64_t k = 1; k < m; k++) {
      /* splitInverse (fft/plan.nim:484-499) with w = (sr[k], -si[k]) */
      T a_r = re[k], a_i = im[k], b_r = re[m - k], b_i = im[m - k];
      T wr = p->spr[k], wi = -p->spi[k];
      T er = a_r + b_r, ei = a_i - b_i, dr = a_r - b_r, di = a_i + b_i;
      T orr = dr * wr - di * wi;
      T oi = dr * wi + di * wr;
      ai[k] = (er - oi) * scale;
      ar[k] = (ei + orr) * scale;
    }
  }
  int inB = SFX(complex_forward)(p, ar, ai, br, bi);
  T *zr = inB ? bi : ai, *zi = inB ? br : ar;
  if (p->n % 2 == 1) {
    for (int64_t i = 0; i < m; i++) x[i] = zr[i];
  } else {
    for (int64_t k = 0; k < m; k++) { x[2 * k] = zr[k]; x[2 * k + 1] = zi[k]; }
  }
}

int SFX(or_rfft)(const T *x, int64_t n, T *re, T *im) {
  SFX(fftplan) p;
  if (SFX(fft_init)(&p, n)) return -1;
  T *work = (T *)malloc(sizeof(T) * (size_t)(4 * p.m));
  SFX(fft_forward)(&p, x, re, im, work);
  free(work);
  SFX(fft_free)(&p);
  return 0;
}

int SFX(or_irfft)(const T *re, const T *im, int64_t n, T *x) {
  SFX(fftplan) p;
  if (SFX(fft_init)(&p, n)) return -1;
  T *work = (T *)malloc(sizeof(T) * (size_t)(4 * p.m));
  SFX(fft_inverse)(&p, re, im, x, work);
  free(work);
  SFX(fft_free)(&p);
  return 0;
}

/* ----------------------------------------------------------------- sums -- */

/* sum64 (core/sums.nim:12-23): eight float64 accumulators. */
static double SFX(sum64)(const T *x, int64_t len) {
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0}, r = 0;
  int64_t n8 = len / 8 * 8;
  for (int64_t i = 0; i < n8; i += 8)
    for (int j = 0; j < 8; j++) acc[j] += (double)x[i + j];
  for (int j = 0; j < 8; j++) r += acc[j];
  for (int64_t k = n8; k < len; k++) r += (double)x[k];
  return r;
}

/* sumIndexed64 (core/sums.nim:25-40). */
static double SFX(sum_indexed64)(const T *x, int64_t len) {
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0}, idx[8], r = 0;
  for (int j = 0; j < 8; j++) idx[j] = (double)j;
  int64_t n8 = len / 8 * 8;
  for (int64_t i = 0; i < n8; i += 8)
    for (int j = 0; j < 8; j++) { acc[j] += idx[j] * (double)x[i + j]; idx[j] += 8; }
  for (int j = 0; j < 8; j++) r += acc[j];
  for (int64_t k = n8; k < len; k++) r += (double)k * (double)x[k];
  return r;
}

/* sumSq64 (gain/stats.nim:82-95). */
static double SFX(sum_sq64)(const T *x, int64_t len) {
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0}, r = 0;
  int64_t n8 = len / 8 * 8;
  for (int64_t i = 0; i < n8; i += 8)
    for (int j = 0; j < 8; j++) { double v = (double)x[i + j]; acc[j] += v * v; }
  for (int j = 0; j < 8; j++) r += acc[j];
  for (int64_t k = n8; k < len; k++) r += (double)x[k] * (double)x[k];
  return r;
}

/* maxAbs (gain/stats.nim:97-117): integer maximum of the sign-cleared bits. */
static double SFX(max_abs)(const T *x, int64_t len) {
  UINT_T m = 0;
  for (int64_t i = 0; i < len; i++) {
    UINT_T b;
    memcpy(&b, &x[i], sizeof(b));
    b &= ABS_MASK;
    if ((SINT_T)b > (SINT_T)m) m = b;
  }
  T v;
  memcpy(&v, &m, sizeof(v));
  return (double)v;
}

/* --------------------------------------------------------- basic kernels -- */

/* demean (core/basic.nim:52-59). */
static void SFX(demean)(T *x, int64_t n) {
  if (n == 0) return;
  T m = (T)(SFX(sum64)(x, n) / (double)n);
  for (int64_t i = 0; i < n; i++) x[i] = x[i] - m;
}

/* detrend (core/basic.nim:61-81). */
static void SFX(detrend)(T *x, int64_t n) {
  if (n == 0) return;
  double nf = (double)n, c = (nf - 1) / 2, sx = SFX(sum64)(x, n), mean = sx / nf;
  double stt = nf * (nf * nf - 1) / 12;
  double slope = n > 1 ? (SFX(sum_indexed64)(x, n) - c * sx) / stt : 0.0;
  T a = (T)(mean - slope * c), b = (T)slope;
  for (int32_t i = 0; i < (int32_t)n; i++) x[i] = x[i] - (a + b * (T)i);
}

/* scale (core/basic.nim:83-90). */
static void SFX(scale)(T *x, int64_t n, double factor) {
  T k = (T)factor;
  for (int64_t i = 0; i < n; i++) x[i] = x[i] * k;
}

/* clip (core/basic.nim:92-103). */
static void SFX(clip)(T *x, int64_t n, double lo, double hi) {
  T l = (T)lo, h = (T)hi;
  for (int64_t i = 0; i < n; i++) {
    T v = x[i];
    x[i] = v > h ? h : (v < l ? l : v);
  }
}

static void SFX(polarity)(T *x, int64_t n) {
  for (int64_t i = 0; i < n; i++) x[i] = -x[i];
}

int SFX(or_demean)(T *x, int64_t n) { SFX(demean)(x, n); return 0; }

/* ------------------------------------------------------------ FFT filter -- */

/* FilterPlan (filter/apply.nim:94-108). */
typedef struct SFX(filterplan) {
  or_spec spec;
  SFX(fftplan) fft;
  T *gain, *fall;
  int64_t pad, margin, samples;
  int reflect, trend;
  double dt, dc_gain;
} SFX(filterplan);

static void SFX(filter_free)(SFX(filterplan) *f) {
  SFX(fft_free)(&f->fft);
  free(f->gain); free(f->fall);
  memset(f, 0, sizeof(*f));
}

/* initFilterPlan (filter/apply.nim:118-151). */
static int SFX(filter_init)(SFX(filterplan) *f, const or_spec *spec, double dt,
                            int64_t samples, int trend, char *err, int errlen) {
  memset(f, 0, sizeof(*f));
  or_filter_info info;
  if (or_filter_info_of(spec, dt, samples, trend, &info, err, errlen)) return -1;
  f->spec = *spec;
  if (SFX(fft_init)(&f->fft, info.nfft)) { or_seterr(err, errlen, "bad nfft"); return -1; }
  int64_t bins = info.nfft / 2 + 1;
  f->gain = (T *)malloc(sizeof(T) * (size_t)bins);
  double df = 1000 / ((double)info.nfft * dt);          /* responseInto (design.nim:257-265) */
  for (int64_t k = 0; k < bins; k++) f->gain[k] = (T)or_response(spec, (double)k * df, dt);
  if (spec->zero_dc && bins > 0) f->gain[0] = (T)0;
  f->fall = (T *)malloc(sizeof(T) * (size_t)(info.pad + 1));
  for (int64_t k = 1; k <= info.pad; k++) {
    double c = cos(OR_PI / 2 * (double)k / (double)(info.pad + 1));
    f->fall[k - 1] = (T)(c * c);
  }
  f->pad = info.pad;
  f->margin = info.margin;
  f->samples = samples;
  f->dt = dt;
  f->dc_gain = info.dc_gain;
  f->trend = trend;
  f->reflect = info.reflect;
  return 0;
}

/* trendOf (filter/edges.nim:16-30). */
static void SFX(trend_of)(int trend, const T *x, int64_t n, double *a, double *b) {
  if (trend == 0) { *a = 0; *b = 0; return; }
  if (trend == 1) { *a = SFX(sum64)(x, n) / (double)n; *b = 0; return; }
  double sx = SFX(sum64)(x, n), nf = (double)n, mean = sx / nf;
  if (n == 1) { *a = mean; *b = 0; return; }
  double c = (nf - 1) / 2;
  double slope = (SFX(sum_indexed64)(x, n) - c * sx) / (nf * (nf * nf - 1) / 12);
  *a = mean - slope * c;
  *b = slope;
}

/* packInto (filter/apply.nim:194-216). */
static void SFX(pack)(const SFX(filterplan) *f, const T *x, int64_t n, double a, double b,
                      T *dst) {
  int64_t p = f->pad, nf = f->fft.n;
  for (int64_t i = 0; i < n; i++) dst[p + i] = (T)((double)x[i] - (a + b * (double)i));
  if (f->reflect) {
    T d0 = dst[p], dn = dst[p + n - 1];
    for (int64_t k = 1; k <= p; k++) {
      T r = dst[p + (k < n - 1 ? k : n - 1)];
      dst[p - k] = (d0 + d0 - r) * f->fall[k - 1];
    }
    for (int64_t k = 1; k <= p; k++) {
      T r = dst[p + n - 1 - (k < n - 1 ? k : n - 1)];
      dst[p + n - 1 + k] = (dn + dn - r) * f->fall[k - 1];
    }
  } else {
    for (int64_t k = 1; k <= p; k++) { dst[p - k] = (T)0; dst[p + n - 1 + k] = (T)0; }
  }
  for (int64_t j = 2 * p + n; j < nf; j++) dst[j] = (T)0;
}

/* unpackFrom (filter/apply.nim:218-231). */
static void SFX(unpack)(const SFX(filterplan) *f, const T *src, double a, double b, T *x,
                        int64_t n) {
  int64_t p = f->pad;
  double gb = b * f->dc_gain;
  double ga = f->spec.zero_dc ? -gb * (double)(n - 1) / 2 : a * f->dc_gain;
  if (ga == 0 && gb == 0) {
    for (int64_t i = 0; i < n; i++) x[i] = src[p + i];
  } else {
    for (int64_t i = 0; i < n; i++) x[i] = (T)((double)src[p + i] + (ga + gb * (double)i));
  }
}

/* single-trace applyFilter (filter/apply.nim:340-363); scratch holds
 * nfft + 2 bins + 4 m values. */
static void SFX(filter_apply)(const SFX(filterplan) *f, T *x, int64_t n, T *scratch) {
  if (n == 0) return;
  int64_t nf = f->fft.n, nb = nf / 2 + 1;
  T *buf = scratch, *re = scratch + nf, *im = scratch + nf + nb, *work = scratch + nf + 2 * nb;
  double a, b;
  SFX(trend_of)(f->trend, x, n, &a, &b);
  SFX(pack)(f, x, n, a, b, buf);
  SFX(fft_forward)(&f->fft, buf, re, im, work);
  for (int64_t k = 0; k < nb; k++) { T g = f->gain[k]; re[k] = re[k] * g; im[k] = im[k] * g; }
  SFX(fft_inverse)(&f->fft, re, im, buf, work);
  SFX(unpack)(f, buf, a, b, x, n);
}

static int64_t SFX(filter_scratch)(const SFX(filterplan) *f) {
  int64_t nf = f->fft.n;
  return nf + 2 * (nf / 2 + 1) + 4 * f->fft.m;
}

int SFX(or_filter)(T *x, int64_t n, const or_spec *spec, double dt_ms, int64_t samples,
                   int32_t trend, char *err, int errlen) {
  SFX(filterplan) f;
  if (SFX(filter_init)(&f, spec, dt_ms, samples, trend, err, errlen)) return -1;
  if (n > samples) { or_seterr(err, errlen, "a trace longer than the plan's samples");
    SFX(filter_free)(&f); return -1; }
  T *s = (T *)malloc(sizeof(T) * (size_t)SFX(filter_scratch)(&f));
  SFX(filter_apply)(&f, x, n, s);
  free(s);
  SFX(filter_free)(&f);
  return 0;
}

/* ------------------------------------------------------------------- AGC -- */

/* fitsIn (gain/gain.nim:112-118). */
static int SFX(fits_in)(double x) {
  if (!or_finite_positive(x)) return 0;
  T t = (T)x;
  return or_finite_positive((double)t);
}

/* blockSums (gain/gain.nim:351-390). */
static void SFX(block_sums)(const T *x, int64_t n, int64_t w, int mode, double *s, double *p) {
#define TERM(v) (mode == 0 ? (v) * (v) : fabs(v))
  int64_t bs = 0;
  while (bs + 2 * w <= n) {
    int64_t mid = bs + w;
    double fwd0 = 0, bwd0 = 0, fwd1 = 0, bwd1 = 0;
    p[bs] = 0;
    for (int64_t k = 0; k < w; k++) {
      double v;
      v = (double)x[bs + k]; fwd0 += TERM(v); p[bs + k + 1] = fwd0;
      v = (double)x[mid - 1 - k]; bwd0 += TERM(v); s[mid - 1 - k] = bwd0;
      v = (double)x[mid + k]; fwd1 += TERM(v); p[mid + k + 1] = fwd1;
      v = (double)x[mid + w - 1 - k]; bwd1 += TERM(v); s[mid + w - 1 - k] = bwd1;
    }
    p[mid] = 0;
    bs += 2 * w;
  }
  while (bs < n) {
    int64_t be = bs + w < n ? bs + w : n;
    double fwd = 0, bwd = 0;
    p[bs] = 0;
    for (int64_t k = 0; k < be - bs; k++) {
      double v;
      v = (double)x[bs + k]; fwd += TERM(v); p[bs + k + 1] = fwd;
      v = (double)x[be - 1 - k]; bwd += TERM(v); s[be - 1 - k] = bwd;
    }
    bs = be;
  }
  if (n % w == 0) p[n] = 0;
#undef TERM
}

typedef struct SFX(agcplan) {
  int mode;
  int64_t half, samples;
  double dt, target, epsilon;
  double *inv, *rev;
} SFX(agcplan);

static void SFX(agc_free)(SFX(agcplan) *a) { free(a->inv); free(a->rev); memset(a, 0, sizeof(*a)); }

/* initAgcPlan (gain/gain.nim:272-302). */
static int SFX(agc_init)(SFX(agcplan) *a, double window_ms, double dt, int64_t samples,
                         int mode, double target, double epsilon, char *err, int errlen) {
  memset(a, 0, sizeof(*a));
  if (!or_finite_positive(dt)) return or_seterr(err, errlen, "dtMs must be finite and positive");
  if (samples < 1 || samples > 2147483647)
    return or_seterr(err, errlen, "an AGC plan needs samples in 1 .. int32.high");
  if (!or_finite_positive(window_ms))
    return or_seterr(err, errlen, "an AGC window must be finite and positive");
  double h = round(window_ms / (2 * dt));
  if (h < 1) return or_seterr(err, errlen, "an AGC window is under 3 samples");
  if (h > (double)(1 << 26)) return or_seterr(err, errlen, "an AGC window is too long");
  if (!or_finite_positive(target)) return or_seterr(err, errlen, "an AGC target must be finite and positive");
  if (!or_finite_positive(epsilon)) return or_seterr(err, errlen, "an AGC epsilon must be finite and positive");
  a->mode = mode;
  a->half = (int64_t)h;
  a->samples = samples;
  a->dt = dt;
  a->target = target;
  a->epsilon = epsilon;
  int64_t H = a->half;
  a->inv = (double *)malloc(sizeof(double) * (size_t)(2 * H + 2));
  a->inv[0] = 0;
  for (int64_t c = 1; c < 2 * H + 2; c++) a->inv[c] = 1 / (double)c;
  a->rev = (double *)malloc(sizeof(double) * (size_t)(2 * H + 1));
  for (int64_t m = 0; m < 2 * H + 1; m++) a->rev[m] = 1 / (double)(2 * H + 1 - m);
  return 0;
}

/* agcImpl (gain/gain.nim:404-461); scratch holds 2n+1 float64. */
static void SFX(agc_apply)(const SFX(agcplan) *a, T *x, int64_t n, double *buf) {
  if (n == 0) return;
  double rms = n == 0 ? 0 : sqrt(SFX(sum_sq64)(x, n) / (double)n);
  if (rms == 0) return;
  int64_t h = a->half, w = 2 * h + 1;
  double *s = buf, *p = buf + n;
  SFX(block_sums)(x, n, w, a->mode, s, p);
  if (!SFX(fits_in)(a->target / rms / a->epsilon)) return;
  double unit = a->mode == 0 ? 1 / (rms * rms) : 1 / rms;
  T k = (T)(a->target / rms);
  T eps = (T)a->epsilon;
  int64_t first = h < n ? h : n;
  int64_t last = first > n - h ? first : n - h;
  int64_t headCut = n - h;
  if (headCut < 0) headCut = 0;
  if (headCut > first) headCut = first;
  int64_t lastBlock = (n - 1) / w * w;
  int64_t tailCut = lastBlock + h;
  if (tailCut < last) tailCut = last;
  if (tailCut > n) tailCut = n;
  double pn = p[n];
  const double *inv = a->inv, *rev = a->rev;
  double unitW = unit / (double)w;
  double headWhole = headCut < first ? pn * unit * inv[n] : 0.0;
#define LEVEL(mean) (a->mode == 0 ? SQRT_T((T)(mean)) : (T)(mean))
#define AGC_RANGE(lo_, hi_, MEAN) \
  for (int64_t i = (lo_); i < (hi_); i++) { \
    T gi = k / (LEVEL(MEAN) + eps); \
    x[i] = (T)(x[i] * gi); \
  }
  AGC_RANGE(0, headCut, p[i + h + 1] * (unit * inv[i + h + 1]))
  AGC_RANGE(headCut, first, headWhole)
  AGC_RANGE(first, last, (s[i - h] + p[i + h + 1]) * unitW)
  AGC_RANGE(last, tailCut, (s[i - h] + pn) * (unit * rev[i + h + 1 - n]))
  AGC_RANGE(tailCut, n, s[i - h] * (unit * rev[i + h + 1 - n]))
#undef AGC_RANGE
#undef LEVEL
}

int SFX(or_agc)(T *x, int64_t n, double window_ms, double dt_ms, int32_t mode, double target,
                double epsilon) {
  SFX(agcplan) a;
  char err[256];
  if (SFX(agc_init)(&a, window_ms, dt_ms, n > 0 ? n : 1, mode, target, epsilon, err,
                    (int)sizeof(err)))
    return -1;
  double *buf = (double *)malloc(sizeof(double) * (size_t)(2 * n + 1));
  SFX(agc_apply)(&a, x, n, buf);
  free(buf);
  SFX(agc_free)(&a);
  return 0;
}

/* ------------------------------------------------------------ statistics -- */

/* percentileAbs (gain/stats.nim:159-182); scratch holds n float64. */
static double SFX(percentile_abs)(const T *x, int64_t n, double p, double *a) {
  if (n == 0) return 0;
  if (isnan(SFX(max_abs)(x, n))) return NAN;
  for (int64_t i = 0; i < n; i++) a[i] = fabs((double)x[i]);
  double h = (double)(n - 1) * p / 100;
  int64_t k = (int64_t)floor(h);
  if (k > n - 1) k = n - 1;
  or_select_k(a, 0, n - 1, k);
  double lower = a[k];
  if (k + 1 >= n) return lower;
  double upper = INFINITY;
  for (int64_t i = k + 1; i < n; i++) upper = a[i] < upper ? a[i] : upper;
  double f = h - (double)k;
  return f == 0 ? lower : lower + f * (upper - lower);
}

double SFX(or_percentile_abs)(const T *x, int64_t n, double p) {
  double *a = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  double r = SFX(percentile_abs)(x, n, p, a);
  free(a);
  return r;
}

/* panelStats (gain/stats.nim:253-304) over the live rows of a block. */
static void SFX(panel_stats)(const T *x, int32_t rows, int32_t cols, const int32_t *row_len,
                             double clip_p, double *out) {
  double sum = 0, mx = 0;
  int64_t used_rows = 0, samples = 0;
  for (int32_t r = 0; r < rows; r++) {
    if (row_len[r] < 0) continue;
    int64_t n = row_len[r];
    const T *row = x + (int64_t)r * cols;
    double s = SFX(sum_sq64)(row, n), m = SFX(max_abs)(row, n);
    if (isnan(m) || isinf(m) || s == INFINITY) continue;
    sum += s;
    mx = m > mx ? m : mx;
    samples += n;
    used_rows++;
  }
  out[0] = 0; out[1] = 0; out[2] = NAN; out[3] = (double)used_rows; out[4] = (double)samples;
  if (samples == 0) return;
  out[0] = sqrt(sum / (double)samples);
  out[1] = mx;
  if (isnan(clip_p)) return;
  if (clip_p == 100) { out[2] = mx; return; }
  uint32_t *hist = (uint32_t *)calloc(OR_HIST_BINS, sizeof(uint32_t));
  for (int32_t r = 0; r < rows; r++) {
    if (row_len[r] < 0) continue;
    int64_t n = row_len[r];
    const T *row = x + (int64_t)r * cols;
    double s = SFX(sum_sq64)(row, n), m = SFX(max_abs)(row, n);
    if (isnan(m) || isinf(m) || s == INFINITY) continue;
    for (int64_t i = 0; i < n; i++) hist[or_key_of((float)fabs((double)row[i]))]++;
  }
  double rank = (double)(samples - 1) * clip_p / 100;
  int64_t k = (int64_t)floor(rank);
  if (k > samples - 1) k = samples - 1;
  double lower = or_order_stat(hist, k, mx);
  double f = rank - (double)k;
  out[2] = (f == 0 || k + 1 >= samples) ? lower
                                         : lower + f * (or_order_stat(hist, k + 1, mx) - lower);
  free(hist);
}

int SFX(or_panel_stats)(const T *x, int32_t rows, int32_t cols, const int32_t *row_len,
                        double clip_percentile, double *out) {
  SFX(panel_stats)(x, rows, cols, row_len, clip_percentile, out);
  return 0;
}

/* ---------------------------------------------------------------- mutes -- */

/* muteTop (core/mute.nim:136-156): x[j] for j >= start_col untouched,
 * start_col - taper <= j < start_col scaled by sin²(π/2 (j - start_col +
 * taper) / taper) in float64, earlier columns zeroed; a NaN start leaves
 * the row alone, an infinite taper is none. */
void SFX(or_mute_top)(T *x, int64_t n, double start_col, double taper_cols) {
  if (isnan(start_col)) return;
  const double taper = taper_cols < INFINITY ? taper_cols : 0.0;
  const double ramp_start = start_col - taper;
  const int64_t zero_end = or_first_at_or_after(ramp_start, n);
  for (int64_t j = 0; j < zero_end; j++) x[j] = (T)0;
  if (taper > 0) {
    const int64_t e = or_first_at_or_after(start_col, n);
    for (int64_t j = zero_end; j < e; j++) {
      const double s = sin(M_PI / 2 * ((double)j - ramp_start) / taper);
      x[j] = (T)((double)x[j] * (s * s));
    }
  }
}

/* muteBottom (core/mute.nim:158-180), the mirror image. */
void SFX(or_mute_bottom)(T *x, int64_t n, double end_col, double taper_cols) {
  if (isnan(end_col)) return;
  const double taper = taper_cols < INFINITY ? taper_cols : 0.0;
  const double ramp_end = end_col + taper;
  const int64_t zero_start = or_first_after(ramp_end, n);
  if (taper > 0) {
    for (int64_t j = or_first_after(end_col, n); j < zero_start; j++) {
      const double s = sin(M_PI / 2 * (ramp_end - (double)j) / taper);
      x[j] = (T)((double)x[j] * (s * s));
    }
  }
  for (int64_t j = zero_start; j < n; j++) x[j] = (T)0;
}

/* ------------------------------------------------------------- pipeline -- */

/* ---------------------------------------------------------------- IIR -- */

/* applyFilter on one trace (filter/iir.nim:363-391, 418-433): the trend
 * removed, the pads (odd reflections before, and after for zero phase, or
 * zeros after), the passes in float64, the trend's steady-state response
 * G(1)·(a + b·i) − b·G'(1) added back.  buf: lead + n + tail doubles. */
static void SFX(iir_run_row)(const or_iir *p, T *x, int64_t n, double *buf) {
  if (n <= 0) return;
  double a, b;
  SFX(trend_of)(p->trend, x, n, &a, &b);
  const int64_t L = p->lead;
  for (int64_t i = 0; i < n; i++) buf[L + i] = (double)x[i] - (a + b * (double)i);
  if (p->reflect) {
    const double d0 = buf[L];
    for (int64_t k = 1; k <= L; k++) buf[L - k] = 2 * d0 - buf[L + (k < n - 1 ? k : n - 1)];
    const double dn = buf[L + n - 1];
    for (int64_t k = 1; k <= p->tail; k++)
      buf[L + n - 1 + k] = 2 * dn - buf[L + n - 1 - (k < n - 1 ? k : n - 1)];
  } else {
    for (int64_t k = 1; k <= p->tail; k++) buf[L + n - 1 + k] = 0;
  }
  or_iir_run(p, buf, L + n + p->tail);
  const double ga = a * p->dc_gain - b * p->dc_slope, gb = b * p->dc_gain;
  for (int64_t i = 0; i < n; i++) x[i] = (T)(buf[L + i] + (ga + gb * (double)i));
}

void SFX(or_iir_apply)(const or_iir *p, T *x, int64_t n) {
  double *buf = (double *)malloc(sizeof(double) * (size_t)(p->lead + n + p->tail + 1));
  SFX(iir_run_row)(p, x, n, buf);
  free(buf);
}

/* ---------------------------------------------------------- gain curve -- */

/* GainCurvePlan (gain/curves.nim:214-249): T(factorAt(t)) at
 * t = tmin + k·dt for k < len. */
typedef struct SFX(gainplan) {
  or_op curve;
  T *table;
  int64_t len;
  double tmin, dt;
} SFX(gainplan);

static void SFX(gain_free)(SFX(gainplan) *g) {
  free(g->table);
  g->table = NULL;
}

/* initGainCurvePlan (gain/curves.nim:214-249) for traces of `samples`. */
static int SFX(gain_init)(SFX(gainplan) *g, const or_op *op, double dt, int64_t samples,
                          char *msg, int msglen) {
  memset(g, 0, sizeof(*g));
  if (!or_finite_positive(dt)) { snprintf(msg, msglen, "dtMs must be finite and positive, got %g", dt); return -1; }
  if (samples < 1) { snprintf(msg, msglen, "a gain curve plan needs samples >= 1"); return -1; }
  if (or_gain_check(op, msg, msglen)) return -1;
  const double lo = op->gain_tmin_ms;
  const double hi = isnan(op->gain_tmax_ms) ? lo + (double)(samples - 1) * dt : op->gain_tmax_ms;
  if (!(isfinite(lo) && isfinite(hi))) { snprintf(msg, msglen, "the table limits must be finite, got %g .. %g", lo, hi); return -1; }
  if (hi < lo) { snprintf(msg, msglen, "tMaxMs %g is before tMinMs %g", hi, lo); return -1; }
  const double span = floor((hi - lo) / dt + 1e-9);
  if (span >= (double)(1 << 26)) { snprintf(msg, msglen, "a gain table of %g samples is too long", span); return -1; }
  g->len = (int64_t)span + 1;
  g->table = (T *)malloc(sizeof(T) * (size_t)g->len);
  g->curve = *op;
  g->tmin = lo;
  g->dt = dt;
  for (int64_t k = 0; k < g->len; k++) {
    const double t = lo + (double)k * dt;
    const double f = or_gain_factor_at(op, t);
    g->table[k] = (T)f;
    if (isinf((double)g->table[k]) || isnan(f)) {
      /* 0 · Inf would turn muted zeros into NaN */
      snprintf(msg, msglen, "the gain at %g ms (%g) overflows the sample type", t, f);
      SFX(gain_free)(g);
      return -1;
    }
  }
  return 0;
}

/* applyGainCurve's workspace form on one row (gain/curves.nim:380-416):
 * the table on x[j0 ..< j1] (two entries interpolated in T off the grid),
 * T(factorAt) elsewhere, a zero staying zero; a NaN start leaves the row. */
static void SFX(gain_apply)(const SFX(gainplan) *g, T *x, int64_t n, double start) {
  if (n <= 0 || isnan(start)) return;
  const or_gain_span sp = or_gain_span_of(g->len, g->tmin, g->dt, n, start);
  for (int64_t j = 0; j < n; j++) {
    if (j >= sp.j0 && j < sp.j1) continue;
    const T fac = (T)or_gain_factor_at(&g->curve, start + (double)j * g->dt);
    if (x[j] != 0) x[j] = x[j] * fac;
  }
  if (sp.j1 > sp.j0) {
    const T *tab = g->table + sp.s;
    if (sp.on_grid) {
      for (int64_t j = sp.j0; j < sp.j1; j++) x[j] = x[j] * tab[j];
    } else {
      const T f = (T)sp.f;
      for (int64_t j = sp.j0; j < sp.j1; j++) {
        const T a = tab[j];
        x[j] = x[j] * (a + f * (tab[j + 1] - a));
      }
    }
  }
}

int SFX(or_gain_curve_apply)(const or_op *op, double dt_ms, int64_t samples, T *x, int64_t n,
                             double start_ms, int64_t *table_len, char *err, int errlen) {
  SFX(gainplan) g;
  char msg[256];
  if (SFX(gain_init)(&g, op, dt_ms, samples, msg, (int)sizeof msg)) return or_seterr(err, errlen, msg);
  if (table_len) *table_len = g.len;
  if (n > samples) {
    SFX(gain_free)(&g);
    return or_seterr(err, errlen, "a trace longer than the plan's samples");
  }
  SFX(gain_apply)(&g, x, n, start_ms);
  SFX(gain_free)(&g);
  return 0;
}

typedef struct SFX(step) {
  int kind;
  double factor, lo, hi;
  or_op mute;          /* a mute step: its op (curve and taper) */
  SFX(gainplan) gain;
  or_iir iir;
  SFX(filterplan) filter;
  SFX(agcplan) agc;
  int stat;
  double target, percentile;
} SFX(step);

typedef struct SFX(plan) {
  int n_steps;
  SFX(step) steps[OR_MAX_STEPS];
  int n_parts;
  int part_a[OR_MAX_STEPS + 1], part_b[OR_MAX_STEPS + 1]; /* [a, b) */
  double dt;
  int64_t samples, record, margin;
  int64_t scratch;   /* T values for filters; float64 for AGC and percentile */
} SFX(plan);

static void SFX(plan_free)(SFX(plan) *p) {
  for (int i = 0; i < p->n_steps; i++) {
    if (p->steps[i].kind == ST_FILTER) SFX(filter_free)(&p->steps[i].filter);
    if (p->steps[i].kind == ST_AGC) SFX(agc_free)(&p->steps[i].agc);
    if (p->steps[i].kind == ST_GAIN_CURVE) SFX(gain_free)(&p->steps[i].gain);
  }
  p->n_steps = 0;
}

/* compileOp + compile (pipeline.nim:392-523). */
static int SFX(plan_compile)(SFX(plan) *P, const or_op *ops, int32_t n_ops, double dt,
                             int64_t samples, int64_t record, char *err, int errlen) {
  memset(P, 0, sizeof(*P));
  if (!(isfinite(dt) && dt > 0)) return or_seterr(err, errlen, "dtMs must be finite and positive");
  if (samples < 1) return or_seterr(err, errlen, "a pipeline needs samples >= 1");
  if (record != 0 && record < samples) return or_seterr(err, errlen, "recordSamples must be 0 or >= samples");
  if (n_ops > OR_MAX_STEPS) return or_seterr(err, errlen, "too many ops");
  P->dt = dt;
  P->samples = samples;
  P->record = record > samples ? record : samples;
  char msg[256];
  for (int i = 0; i < n_ops; i++) {
    const or_op *op = &ops[i];
    SFX(step) *s = &P->steps[P->n_steps];
    memset(s, 0, sizeof(*s));
    msg[0] = 0;
    int bad = 0;
    switch (op->kind) {
      case 0: s->kind = ST_DEMEAN; break;
      case 1: s->kind = ST_DETREND; break;
      case 5: s->kind = ST_POLARITY; break;
      case 2:
        if (!isfinite(op->value)) { bad = 1; snprintf(msg, sizeof msg, "the factor must be finite"); break; }
        if (!isfinite((double)(T)op->value)) { bad = 1; snprintf(msg, sizeof msg, "the factor overflows the sample type"); break; }
        s->kind = ST_SCALE; s->factor = op->value; break;
      case 3: {
        if (!isfinite(op->value)) { bad = 1; snprintf(msg, sizeof msg, "the gain must be finite"); break; }
        double f = pow(10.0, op->value / 20);
        if (!isfinite(f) || !isfinite((double)(T)f)) { bad = 1; snprintf(msg, sizeof msg, "the gain overflows the sample type"); break; }
        s->kind = ST_SCALE; s->factor = f; break;
      }
      case 4:
        if (isnan(op->lo) || isnan(op->hi) || op->lo > op->hi) { bad = 1; snprintf(msg, sizeof msg, "clip limits must satisfy lo <= hi"); break; }
        s->kind = ST_CLIP; s->lo = op->lo; s->hi = op->hi; break;
      case 6:
      case 7:
        if (or_mute_check(op, msg, (int)sizeof msg)) { bad = 1; break; }
        s->kind = op->kind == 6 ? ST_MUTE_TOP : ST_MUTE_BOTTOM;
        s->mute = *op;
        break;
      case 8:
        s->kind = ST_FILTER;
        if (SFX(filter_init)(&s->filter, &op->spec, dt, samples, op->trend, msg, (int)sizeof msg)) bad = 1;
        break;
      case 10:
        s->kind = ST_AGC;
        if (SFX(agc_init)(&s->agc, op->window_ms, dt, samples, op->mode, op->target, op->epsilon, msg, (int)sizeof msg)) bad = 1;
        break;
      case 11:
        if (!or_finite_positive(op->target)) { bad = 1; snprintf(msg, sizeof msg, "a normalize target must be finite and positive"); break; }
        if (!(op->percentile >= 0 && op->percentile <= 100)) { bad = 1; snprintf(msg, sizeof msg, "a percentile must be in 0 .. 100"); break; }
        s->kind = op->scope == 1 ? ST_NORM_PANEL : ST_NORM_TRACE;
        s->stat = op->stat; s->target = op->target; s->percentile = op->percentile;
        break;
      case 9:
        /* iirOp (pipeline.nim:428-430) */
        if (or_iir_init(&s->iir, &op->spec, dt, samples, op->iir_phase, op->trend, msg, (int)sizeof msg)) { bad = 1; break; }
        s->kind = ST_IIR;
        break;
      case 12:
        /* gainCurveOp: tabulated to the end of the record (pipeline.nim:440-442) */
        if (SFX(gain_init)(&s->gain, op, dt, P->record, msg, (int)sizeof msg)) { bad = 1; break; }
        s->kind = ST_GAIN_CURVE;
        break;
      default:
        bad = 1; snprintf(msg, sizeof msg, "not part of the panel hot path");
    }
    if (bad) {
      char full[320];
      snprintf(full, sizeof full, "op %d (%s): %s", i, or_op_name(op->kind), msg);
      or_seterr(err, errlen, full);
      SFX(plan_free)(P);
      return -1;
    }
    P->n_steps++;
  }
  int start = 0;
  int64_t need = 1;
  for (int i = 0; i < P->n_steps; i++) {
    SFX(step) *s = &P->steps[i];
    if (s->kind == ST_FILTER) {
      P->margin += s->filter.margin;
      int64_t f = SFX(filter_scratch)(&s->filter);
      need = f > need ? f : need;
    }
    if (s->kind == ST_AGC) {
      P->margin += s->agc.half;
      int64_t f = 2 * samples + 1;
      need = f > need ? f : need;
    }
    if (s->kind == ST_NORM_TRACE) need = samples > need ? samples : need;
    if (s->kind == ST_IIR) {
      P->margin += s->iir.pad;
      int64_t f = s->iir.lead + samples + s->iir.tail;
      need = f > need ? f : need;
    }
    if (s->kind == ST_NORM_PANEL && i > start) {
      P->part_a[P->n_parts] = start; P->part_b[P->n_parts] = i; P->n_parts++;
      start = i;
    }
  }
  P->part_a[P->n_parts] = start; P->part_b[P->n_parts] = P->n_steps; P->n_parts++;
  P->scratch = need;
  return 0;
}

/* runStep (pipeline.nim:671-688) on one live row of n samples; key: the
 * row's (t0Ms, offset), first: ChunkMeta.firstSample, dt: the plan's. */
static void SFX(run_step)(const SFX(step) *s, T *x, int64_t n, void *scratch, const double *key,
                          int64_t first, double dt) {
  switch (s->kind) {
    case ST_MUTE_TOP:
    case ST_MUTE_BOTTOM: {
      /* muteCol (core/mute.nim:189-194): the row's column of the mute time */
      const double t0 = key ? key[0] : 0.0, off = key ? key[1] : NAN;
      const double col = (or_mute_time_ms(&s->mute, off) - t0) / dt - (double)first;
      const double taper_cols = s->mute.taper_ms / dt;
      if (s->kind == ST_MUTE_TOP) SFX(or_mute_top)(x, n, col, taper_cols);
      else SFX(or_mute_bottom)(x, n, col, taper_cols);
      break;
    }
    case ST_GAIN_CURVE:
      /* sample 0 of the row is at sampleTimeMs(r, 0) = t0Ms + firstSample·dt
       * (core/types.nim:237-240) */
      SFX(gain_apply)(&s->gain, x, n, (key ? key[0] : 0.0) + (double)first * dt);
      break;
    case ST_IIR: SFX(iir_run_row)(&s->iir, x, n, (double *)scratch); break;
    case ST_DEMEAN: SFX(demean)(x, n); break;
    case ST_DETREND: SFX(detrend)(x, n); break;
    case ST_POLARITY: SFX(polarity)(x, n); break;
    case ST_SCALE: SFX(scale)(x, n, s->factor); break;
    case ST_CLIP: SFX(clip)(x, n, s->lo, s->hi); break;
    case ST_FILTER: SFX(filter_apply)(&s->filter, x, n, (T *)scratch); break;
    case ST_AGC: SFX(agc_apply)(&s->agc, x, n, (double *)scratch); break;
    case ST_NORM_TRACE: {
      /* normalize (gain/gain.nim:198-219) */
      double level;
      if (s->stat == 0) level = SFX(max_abs)(x, n);
      else if (s->stat == 1) level = n == 0 ? 0 : sqrt(SFX(sum_sq64)(x, n) / (double)n);
      else level = SFX(percentile_abs)(x, n, s->percentile, (double *)scratch);
      if (or_finite_positive(level) && SFX(fits_in)(s->target / level))
        SFX(scale)(x, n, s->target / level);
      break;
    }
    default: break;
  }
}

/* panelFactor (gain/gain.nim:221-240). */
static double SFX(panel_factor)(const SFX(step) *s, const T *x, int32_t rows, int32_t cols,
                                const int32_t *row_len) {
  double st[5];
  SFX(panel_stats)(x, rows, cols, row_len, s->stat == 2 ? s->percentile : NAN, st);
  double level = s->stat == 0 ? st[1] : (s->stat == 1 ? st[0] : st[2]);
  if (or_finite_positive(level) && SFX(fits_in)(s->target / level)) return s->target / level;
  return NAN;
}

/* Rows `r0 ..< r1` of part `part` (runPart, pipeline.nim:690-723), tile by
 * tile of `tile` rows. */
static void SFX(run_part_rows)(const SFX(plan) *P, int part, const T *src, T *dst,
                               int32_t r0, int32_t r1, int32_t cols, const int32_t *row_len,
                               double factor, int64_t tile, void *scratch, const double *keys,
                               int64_t first_sample) {
  int first = P->part_a[part];
  if (P->part_b[part] > first && P->steps[first].kind == ST_NORM_PANEL) first++;
  for (int32_t a = r0; a < r1; a += (int32_t)tile) {
    int32_t b = a + (int32_t)tile < r1 ? a + (int32_t)tile : r1;
    if (src && src != dst)
      memcpy(dst + (int64_t)a * cols, src + (int64_t)a * cols, sizeof(T) * (size_t)(b - a) * (size_t)cols);
    if (!isnan(factor))
      for (int32_t r = a; r < b; r++)
        if (row_len[r] >= 0) SFX(scale)(dst + (int64_t)r * cols, row_len[r], factor);
    for (int i = first; i < P->part_b[part]; i++)
      for (int32_t r = a; r < b; r++)
        if (row_len[r] >= 0)
          SFX(run_step)(&P->steps[i], dst + (int64_t)r * cols, row_len[r], scratch,
                        keys ? keys + 2 * (int64_t)r : NULL, first_sample, P->dt);
  }
}

static int64_t SFX(tile_rows)(int32_t cols) {
  /* tileRows (pipeline.nim:612-617) */
  int64_t c = cols > 0 ? cols : 1;
  int64_t r = (256 * 1024) / (c * (int64_t)sizeof(T));
  int64_t t = r / 8 * 8;
  return t > 8 ? t : 8;
}

static int SFX(check_apply)(const SFX(plan) *P, int32_t cols, char *err, int errlen) {
  if (cols > P->samples) return or_seterr(err, errlen, "the block has more columns than the plan takes");
  return 0;
}

int SFX(or_pipeline_apply_keys)(const or_op *ops, int32_t n_ops, double dt_ms, int64_t samples,
                                int64_t record_samples, const T *src, T *dst, int32_t rows,
                                int32_t cols, const int32_t *row_len, const double *keys,
                                int64_t first, char *err, int errlen) {
  SFX(plan) *P = (SFX(plan) *)malloc(sizeof(SFX(plan)));
  if (SFX(plan_compile)(P, ops, n_ops, dt_ms, samples, record_samples, err, errlen)) {
    free(P);
    return -1;
  }
  if (SFX(check_apply)(P, cols, err, errlen)) { SFX(plan_free)(P); free(P); return -1; }
  void *scratch = malloc(sizeof(double) * (size_t)(P->scratch + 8));
  int64_t tile = SFX(tile_rows)(cols);
  for (int part = 0; part < P->n_parts; part++) {
    int from0 = part == 0;
    double factor = NAN;
    int a = P->part_a[part];
    if (P->part_b[part] > a && P->steps[a].kind == ST_NORM_PANEL)
      factor = SFX(panel_factor)(&P->steps[a], from0 ? src : dst, rows, cols, row_len);
    SFX(run_part_rows)(P, part, from0 ? src : NULL, dst, 0, rows, cols, row_len, factor, tile,
                       scratch, keys, first);
  }
  free(scratch);
  SFX(plan_free)(P);
  free(P);
  return 0;
}

int SFX(or_pipeline_apply)(const or_op *ops, int32_t n_ops, double dt_ms, int64_t samples,
                           int64_t record_samples, const T *src, T *dst, int32_t rows,
                           int32_t cols, const int32_t *row_len, char *err, int errlen) {
  return SFX(or_pipeline_apply_keys)(ops, n_ops, dt_ms, samples, record_samples, src, dst, rows,
                                     cols, row_len, NULL, 0, err, errlen);
}

/* ---------------------------------------------------------- correlation -- */

typedef struct SFX(corrplan) {
  SFX(fftplan) fft;
  T *mulr, *muli;
  int64_t shift, record, out, pilot;
  double dt, scale;
} SFX(corrplan);

static void SFX(corr_free)(SFX(corrplan) *c) {
  SFX(fft_free)(&c->fft);
  free(c->mulr); free(c->muli);
  memset(c, 0, sizeof(*c));
}

/* initCorrelationPlan (vibroseis/correlate.nim:82-160). */
static int SFX(corr_init)(SFX(corrplan) *c, const double *pilot, int64_t plen, double pdt,
                          double pt0, double dt, int64_t record, int64_t out_samples,
                          int mode, double eps_percent, int normalize, char *err, int errlen) {
  memset(c, 0, sizeof(*c));
  /* Pilot.validate (vibroseis/pilot.nim:89-106) */
  if (plen < 2) return or_seterr(err, errlen, "a pilot needs at least two samples");
  if (!(pdt > 0 && pdt < INFINITY)) return or_seterr(err, errlen, "a pilot's dtMs must be finite and positive");
  if (!isfinite(pt0)) return or_seterr(err, errlen, "a pilot's t0Ms must be finite");
  double energy = 0;
  for (int64_t i = 0; i < plen; i++) {
    if (!isfinite(pilot[i])) return or_seterr(err, errlen, "a pilot sample is not finite");
    energy += pilot[i] * pilot[i];
  }
  if (energy == 0) return or_seterr(err, errlen, "the pilot is all zeros");
  if (!(dt > 0 && dt < INFINITY)) return or_seterr(err, errlen, "dtMs must be finite and positive");
  if (!(fabs(pdt - dt) <= 1e-9 * dt)) return or_seterr(err, errlen, "the pilot is sampled at another interval");
  if (record < 1) return or_seterr(err, errlen, "recordSamples must be >= 1");
  int64_t outN = out_samples == 0 ? record - plen : out_samples;
  if (outN < 1) return or_seterr(err, errlen, "the pilot is not shorter than the record");
  if (out_samples < 0 || outN > record) return or_seterr(err, errlen, "outSamples must be in 0 .. recordSamples");
  if (mode == 1 && !(eps_percent > 0 && eps_percent < INFINITY)) return or_seterr(err, errlen, "epsPercent must be finite and positive");
  double s = pt0 / dt;
  if (fabs(s) >= (double)(record + plen)) return or_seterr(err, errlen, "the pilot starts too far from the time break");
  int64_t whole = (int64_t)round(s);
  double frac = s - (double)whole;
  int64_t n = or_next_fast_len(record + plen + 64 + (whole < 0 ? -whole : whole), 1);
  if (SFX(fft_init)(&c->fft, n)) return or_seterr(err, errlen, "bad nfft");
  int64_t bins = n / 2 + 1;
  fftplan_f64 pf;
  fft_init_f64(&pf, n);
  double *x = (double *)calloc((size_t)n, sizeof(double));
  double *pr = (double *)malloc(sizeof(double) * (size_t)bins);
  double *pi = (double *)malloc(sizeof(double) * (size_t)bins);
  double *work = (double *)malloc(sizeof(double) * (size_t)(4 * pf.m));
  for (int64_t j = 0; j < plen; j++) x[j] = pilot[j];
  fft_forward_f64(&pf, x, pr, pi, work);
  double peak_pow = 0;
  for (int64_t k = 0; k < bins; k++) {
    double q = pr[k] * pr[k] + pi[k] * pi[k];
    peak_pow = q > peak_pow ? q : peak_pow;
  }
  double eps = mode == 1 ? eps_percent / 100 * peak_pow : 0.0;
  double *w = (double *)malloc(sizeof(double) * (size_t)bins);
  double peak = 0;
  for (int64_t k = 0; k < bins; k++) {
    double p2 = pr[k] * pr[k] + pi[k] * pi[k];
    w[k] = mode == 1 ? 1 / (p2 + eps) : 1.0;
    double g = p2 * w[k];
    peak += (k == 0 || (n % 2 == 0 && k == bins - 1)) ? g : 2 * g;
  }
  peak /= (double)n;
  double scale = normalize ? 1 / peak : 1.0;
  c->mulr = (T *)malloc(sizeof(T) * (size_t)bins);
  c->muli = (T *)malloc(sizeof(T) * (size_t)bins);
  for (int64_t k = 0; k < bins; k++) {
    /* conj(P) * (w scale) * exp(i 2 pi k frac / n), with std/complex's products */
    double ws = w[k] * scale;
    double ar = pr[k] * ws, ai = -pi[k] * ws;
    double th = 2 * OR_PI * (double)k * frac / (double)n;
    double er = exp(0.0) * cos(th), ei = exp(0.0) * sin(th);
    c->mulr[k] = (T)(ar * er - ai * ei);
    c->muli[k] = (T)(ar * ei + ai * er);
  }
  if (n % 2 == 0) c->muli[bins - 1] = (T)0;
  c->shift = whole;
  c->record = record;
  c->out = outN;
  c->pilot = plen;
  c->dt = dt;
  c->scale = scale;
  free(x); free(pr); free(pi); free(work); free(w);
  fft_free_f64(&pf);
  return 0;
}

/* Block correlateInto on one used row (vibroseis/correlate.nim:241-290,
 * 421-464): gather to the live length, forward, multiplier, inverse, lag
 * spans. */
static void SFX(corr_row)(const SFX(corrplan) *c, const T *src, int64_t live, T *dst,
                          int64_t cols, T *scratch) {
  int64_t n = c->fft.n, bins = n / 2 + 1;
  T *buf = scratch, *re = scratch + n, *im = scratch + n + bins, *work = scratch + n + 2 * bins;
  for (int64_t j = 0; j < live; j++) buf[j] = src[j];
  for (int64_t j = live; j < n; j++) buf[j] = (T)0;
  SFX(fft_forward)(&c->fft, buf, re, im, work);
  for (int64_t k = 0; k < bins; k++) {
    T a = re[k], b = im[k], mr = c->mulr[k], mi = c->muli[k];
    re[k] = a * mr - b * mi;
    im[k] = a * mi + b * mr;
  }
  SFX(fft_inverse)(&c->fft, re, im, buf, work);
  /* lagSpans (vibroseis/correlate.nim:265-274) */
  int64_t wrapped = 1 - c->pilot - c->shift;
  wrapped = wrapped < 0 ? 0 : (wrapped > cols ? cols : wrapped);
  int64_t direct = -c->shift;
  direct = direct < wrapped ? wrapped : (direct > cols ? cols : direct);
  int64_t last = c->record - c->shift;
  last = last < direct ? direct : (last > cols ? cols : last);
  int64_t wrap = c->shift + n;
  for (int64_t m = 0; m < wrapped; m++) dst[m] = (T)0;
  for (int64_t m = last; m < cols; m++) dst[m] = (T)0;
  for (int64_t m = wrapped; m < direct; m++) dst[m] = buf[m + wrap];
  for (int64_t m = direct; m < last; m++) dst[m] = buf[m + c->shift];
}

int SFX(or_correlate)(const double *pilot, int64_t pilot_len, double pilot_dt_ms,
                      double pilot_t0_ms, double dt_ms, int64_t record_samples,
                      int64_t out_samples, int32_t mode, double eps_percent, int32_t normalize,
                      const T *src, int32_t rows, int32_t src_cols, const int32_t *row_len,
                      T *dst, int32_t dst_cols, char *err, int errlen) {
  SFX(corrplan) c;
  if (SFX(corr_init)(&c, pilot, pilot_len, pilot_dt_ms, pilot_t0_ms, dt_ms, record_samples,
                     out_samples, mode, eps_percent, normalize, err, errlen))
    return -1;
  /* checkCorrelate (vibroseis/correlate.nim:328-347) */
  if (src_cols > c.record) { SFX(corr_free)(&c); return or_seterr(err, errlen, "the records are wider than the plan's recordSamples"); }
  if (dst_cols != c.out) { SFX(corr_free)(&c); return or_seterr(err, errlen, "dst must be outSamples wide"); }
  int64_t n = c.fft.n;
  T *scratch = (T *)malloc(sizeof(T) * (size_t)(n + 2 * (n / 2 + 1) + 4 * c.fft.m));
  for (int32_t r = 0; r < rows; r++) {
    T *d = dst + (int64_t)r * dst_cols;
    if (row_len[r] < 0) {
      for (int32_t m = 0; m < dst_cols; m++) d[m] = (T)0;
      continue;
    }
    SFX(corr_row)(&c, src + (int64_t)r * src_cols, row_len[r], d, dst_cols, scratch);
  }
  free(scratch);
  SFX(corr_free)(&c);
  return 0;
}

/* --------------------------------------------------------- first breaks -- */

/* liveSpan (stack/stack.nim:292-300): the first and last non-zero sample;
 * *a > *b when all are zero. */
static void SFX(live_span)(const T *x, int64_t n, int64_t *a, int64_t *b) {
  int64_t i = 0;
  while (i < n && x[i] == 0) i++;
  if (i == n) { *a = 0; *b = -1; return; }
  int64_t j = n - 1;
  while (x[j] == 0) j--;
  *a = i;
  *b = j;
}

/* meanPower (qc/snr.nim:151-161) over the columns c0 .. c1 (inclusive) cut
 * to 0 .. n - 1; NaN when none. */
static double SFX(mean_power)(const T *x, int64_t n, int64_t c0, int64_t c1) {
  const int64_t a = c0 > 0 ? c0 : 0, b = c1 < n - 1 ? c1 : n - 1;
  if (a > b) return NAN;
  double s = 0;
  for (int64_t j = a; j <= b; j++) {
    const double v = (double)x[j];
    s += v * v;
  }
  return s / (double)(b - a + 1);
}

/* stretchOf (qc/firstbreak.nim:394-430): what the pick of one row reads. */
#ifndef OR_FB_STRETCH
#define OR_FB_STRETCH
typedef struct { int64_t first, len, lo, hi; double origin; } or_fb_stretch;
#endif

static or_fb_stretch SFX(fb_stretch)(const or_fb_plan *P, const T *src, int64_t n,
                                     double offset, double t0_ms, double start_col) {
  or_fb_stretch s = {0, 0, 0, 0, NAN};
  if (n < 2) return s;
  const or_fb_params *pr = &P->p;
  const double dt = P->dt_ms;
  double t_lo = isnan(pr->search_start_ms) ? -INFINITY : pr->search_start_ms;
  double t_hi = isnan(pr->search_end_ms) ? INFINITY : pr->search_end_ms;
  if (!isnan(pr->guide_velocity)) {
    if (isnan(offset)) return s;
    const double tl = pr->guide_t0_ms + 1000 * fabs(offset) / pr->guide_velocity;
    t_lo = fmax(t_lo, tl - pr->before_ms);
    t_hi = fmin(t_hi, tl + pr->after_ms);
  }
  const double t0c = t0_ms + start_col * dt;
  const double origin = isinf(t0c) ? NAN : t0c;
  if (isnan(origin) && (t_lo > -INFINITY || t_hi < INFINITY)) return s;
  const double c_lo = isnan(origin) ? -INFINITY : (t_lo - origin) / dt;
  const double c_hi = isnan(origin) ? INFINITY : (t_hi - origin) / dt;
  int64_t la, lb;
  SFX(live_span)(src, n, &la, &lb);
  if (lb - la + 1 < 2) return s;
  const int64_t lo_c = c_lo <= 0 ? 0 : c_lo >= (double)n ? n : (int64_t)ceil(c_lo);
  const int64_t hi_c = c_hi < 0 ? 0 : c_hi >= (double)(n - 1) ? n : (int64_t)floor(c_hi) + 1;
  const int64_t lo = la > lo_c ? la : lo_c;
  const int64_t hi = lb + 1 < hi_c ? lb + 1 : hi_c;
  if (lo >= hi) return s;
  const int64_t a = P->whole ? la : (la > lo - P->before ? la : lo - P->before);
  const int64_t b = P->whole ? lb + 1 : (lb + 1 < hi + P->after ? lb + 1 : hi + P->after);
  s.first = a;
  s.len = b - a;
  s.lo = lo - a;
  s.hi = hi - a;
  s.origin = origin;
  return s;
}

/* The pick on the m samples x of a stretch (pickStretch, qc/firstbreak.nim:
 * 432-461, with prefixEnergy and argmaxAttribute :302-376).  p, a: m + 1
 * doubles of scratch. */
static void SFX(fb_pick_stretch)(const or_fb_plan *P, const T *x, or_fb_stretch st, double *p,
                                 double *a, double *time_ms, int64_t *col, double *snr_db,
                                 double *conf) {
  *time_ms = NAN;
  *col = -1;
  *snr_db = NAN;
  *conf = 0;
  const int64_t m = st.len;
  double s = 0;
  p[0] = 0;
  for (int64_t i = 0; i < m; i++) {
    const double v = (double)x[i];
    s += v * v;
    p[i + 1] = s;
  }
  const double total = p[m];
  if (!(total > 0) || total == INFINITY) return;
  const double mean = total / (double)m;
  const double eps = 1e-6 * mean;
  const double beta = P->p.stabilization * mean * (double)P->w;
  const int64_t w = P->w, lta = P->lta, lo = st.lo, hi = st.hi, n = m;
  /* argmaxAttribute */
  int64_t best_i = -1;
  double best = 0.0;
  const int64_t sw = w / 2 > 1 ? w / 2 : 1;
  const int64_t last_read = n - 1 < hi - 2 + sw ? n - 1 : hi - 2 + sw;
  int64_t i0 = 0;
  if (P->p.method == 0) {
    i0 = lo > w + sw ? lo : w + sw;
    for (int64_t i = i0 - 1; i <= last_read; i++) {
      const int64_t s0 = i + 1 - w > 0 ? i + 1 - w : 0;
      const int64_t m0 = s0 - lta > 0 ? s0 - lta : 0;
      a[i] = s0 - m0 < sw ? 0.0
                          : ((p[i + 1] - p[s0]) / (double)(i + 1 - s0)) /
                                ((p[s0] - p[m0]) / (double)(s0 - m0) + eps);
    }
  } else if (P->p.method == 1) {
    const int64_t q = w / 4 > 1 ? w / 4 : 1;
    for (int64_t i = lo; i < hi; i++) {
      int64_t k = i < n - i ? i : n - i;
      k = w < k ? w : k;
      if (k >= sw) {
        const int64_t e = n < i + q ? n : i + q;
        const double r = (p[i + k] - p[i]) / ((p[i] - p[i - k]) + eps * (double)k) * sqrt(p[e] - p[i]);
        if (r > best) {
          best = r;
          best_i = i;
        }
      }
    }
  } else {
    i0 = lo > w ? lo : w;
    for (int64_t i = i0 - 1; i <= last_read; i++) {
      const int64_t s0 = i + 1 - w > 0 ? i + 1 - w : 0;
      a[i] = (p[i + 1] - p[s0]) / (p[i + 1] + beta + eps);
    }
  }
  if (P->p.method != 1) {
    /* risePick: the largest mean rise of a[] over sw steps */
    for (int64_t i = i0; i < hi; i++) {
      const int64_t j = n - 1 < i - 1 + sw ? n - 1 : i - 1 + sw;
      if (j > i - 1) {
        const double r = (a[j] - a[i - 1]) / (double)(j - i + 1);
        if (r > best) {
          best = r;
          best_i = i;
        }
      }
    }
  }
  if (best_i < 0) return;
  const int64_t c = best_i;
  const int64_t gap_end = c - w / 2 > 0 ? c - w / 2 : 0;
  *col = st.first + c;
  *time_ms = st.origin + (double)(st.first + c) * P->dt_ms;
  const int64_t nz0 = gap_end - P->snr > 0 ? gap_end - P->snr : 0;
  const double noise = SFX(mean_power)(x, m, nz0, gap_end - 1);
  if (noise > 0) {
    const int64_t sig1 = m < c + P->snr ? m : c + P->snr;
    const double sg = SFX(mean_power)(x, m, c, sig1 - 1);
    /* powerSnrDb (qc/snr.nim:163-177), seRatio, floor -Inf */
    const double ss = sg + eps, nn = noise + eps;
    double snr = NAN;
    if (!isnan(ss) && nn > 0) {
      const double r = ss / nn;
      snr = isnan(r) ? NAN : (r > 0 ? 10 * log10(r) : -INFINITY);
    }
    if (!isnan(snr)) {
      *snr_db = snr;
      double cf = (snr - P->p.low_db) / (P->p.high_db - P->p.low_db);
      *conf = cf < 0 ? 0.0 : (cf > 1 ? 1.0 : cf);
    }
  }
}

void SFX(or_fb_pick)(const or_fb_plan *P, const T *src, int32_t rows, int32_t cols,
                     const int32_t *row_len, const double *keys, int64_t first_sample,
                     double *time_ms, int64_t *col, double *snr_db, double *conf) {
  const int64_t cap = P->samples + 1;
  double *p = (double *)malloc(sizeof(double) * (size_t)cap);
  double *a = (double *)malloc(sizeof(double) * (size_t)cap);
  T *copy = (T *)malloc(sizeof(T) * (size_t)(cols > 0 ? cols : 1));
  for (int32_t r = 0; r < rows; r++) {
    time_ms[r] = NAN;
    col[r] = -1;
    snr_db[r] = NAN;
    conf[r] = 0;
    if (row_len[r] < 0) continue;
    const int64_t n = row_len[r] < cols ? row_len[r] : cols;
    const T *x = src + (int64_t)r * cols;
    const double t0 = keys ? keys[2 * r] : 0.0, off = keys ? keys[2 * r + 1] : NAN;
    or_fb_stretch st = SFX(fb_stretch)(P, x, n, off, t0, (double)first_sample);
    if (st.len == 0) continue;
    const T *y = x + st.first;
    if (P->filtered) {
      for (int64_t j = 0; j < st.len; j++) copy[j] = y[j];
      SFX(or_iir_apply)(&P->filter, copy, st.len);
      y = copy;
    }
    SFX(fb_pick_stretch)(P, y, st, p, a, time_ms + r, col + r, snr_db + r, conf + r);
  }
  free(p);
  free(a);
  free(copy);
}
