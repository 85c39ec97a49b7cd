// Device building blocks of the gfx950 panel kernels, templated on the
// threads that share one row (NTH, a multiple of 64).
//
// The real FFT of n points is a complex Stockham FFT of m = n/2 points on the
// interleaved samples (fft/plan.nim:501-594); its stages ping-pong between two
// LDS buffers (m complex values, plus one spare per first-pass butterfly),
// two stages per pass where the fused butterfly holds at most 16 points.  Split, gain and the inverse's unsplit
// are one fused pass over the bins, and the inverse transform is the forward
// one on swapped (im, re) pairs, so the swap back is an XOR of the sample
// index.  Sums (demean, trend, rms, AGC windows) run in float64 as in the
// reference, reduced in a fixed order (a row's bits depend on the row and
// its plan only); the AGC window sums are block prefix and suffix sums
// without subtraction.
#pragma once
#include "pseg_internal.hpp"

#include <math.h>

#include <type_traits>

namespace pseg {

namespace {


// Plan tables live in global memory; loading them through global (not
// flat) pointers keeps them off the LDS path.
template <typename X>
using gptr = const X __attribute__((address_space(1))) *;
template <typename X>
__device__ __forceinline__ gptr<X> as_global(const void *p) {
  return (gptr<X>)(const X *)p;
}
// A row of a panel in global memory, written: stores through it are global
// (not flat) stores, with the row's base in SGPRs and a 32-bit lane offset.
template <typename X>
using gmut = X __attribute__((address_space(1))) *;
template <typename X>
__device__ __forceinline__ gmut<X> as_global_mut(void *p) {
  return (gmut<X>)(X *)p;
}

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

// threadIdx.x behind an empty asm barrier: values computed from it inside a
// step cannot be hoisted out of the loop over the plan's steps (which would
// keep a dozen per-sample indices live across every step and spill them).
__device__ __forceinline__ int opaque_tid() {
  int t = threadIdx.x;
  asm volatile("" : "+v"(t));
  return t;
}

// A per-lane index behind an empty asm barrier: element offsets added to it
// stay uniform (SGPR) adds instead of being folded into a per-lane multiply.
__device__ __forceinline__ int opaque_v(int i) {
  asm volatile("" : "+v"(i));
  return i;
}

// A uniform value behind an empty asm barrier (kept in an SGPR): what the
// steps compute from the row length (trend denominators, run lengths) is
// computed again in each step instead of hoisted out of the loop over the
// steps, where it would stay live across every step and spill.
__device__ __forceinline__ int opaque_s(int v) {
  asm volatile("" : "+s"(v));
  return v;
}

// Plan parameters are the same on every lane; reading them through
// readfirstlane puts them (and everything computed from them) in SGPRs
// instead of a VGPR per lane.
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t uni(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}
__device__ __forceinline__ double uni(double v) {
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xFFFFFFFF));
  int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <typename P>
__device__ __forceinline__ P *uni_ptr(P *p) {
  unsigned long long b = (unsigned long long)p;
  unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(b & 0xFFFFFFFFull));
  unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(b >> 32));
  return (P *)(((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ FastDiv uni(FastDiv d) {
  FastDiv r;
  r.d = uni(d.d);
  r.mul = uni(d.mul);
  r.shift = uni(d.shift);
  return r;
}

// Ends one iteration of an unrolled per-sample loop: the scheduler may not
// interleave iterations, so a heavy loop body holds one sample's temporaries
// at a time instead of PER of them.
#define PSEG_ONE_AT_A_TIME() __builtin_amdgcn_sched_barrier(0)

__device__ __forceinline__ uint32_t fdiv(uint32_t n, FastDiv d) {
  return (__umulhi(n, d.mul) + n) >> d.shift;
}

// --------------------------------------------------------------- reductions

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// LDS for block reductions: two slots used in turn.  A reduction writes its
// slot, waits at one barrier and reads it; the next one uses the other slot,
// so none needs a barrier before its write (every wave has finished reading
// a slot before any wave passes the barrier of the reduction after it).
// Every thread takes the slots in the same order.
struct RedSlots {
  double *base;  // 2 * size doubles
  int size;
  int flip;
  __device__ __forceinline__ double *next() {
    double *p = base + (flip ? size : 0);
    flip ^= 1;
    return p;
  }
};

// The sum of v over the block, the same value on every thread, in a fixed
// order (lane 0 of each wave, then the waves in order): a row's result does
// not depend on anything but the row.  `red`: a slot no wave still reads.
template <int NTH>
__device__ __forceinline__ double block_sum(double v, double *red) {
  const int tid_ = opaque_tid();
  v = wave_sum(v);
  if ((tid_ & 63) == 0) red[tid_ >> 6] = v;
  __syncthreads();
  double t = 0;
#pragma unroll
  for (int i = 0; i < NTH / 64; i++) t += red[i];
  return t;
}

template <int NTH>
__device__ __forceinline__ void block_sum2(double &a, double &b, double *red) {
  const int tid_ = opaque_tid();
  a = wave_sum(a);
  b = wave_sum(b);
  if ((tid_ & 63) == 0) {
    red[tid_ >> 6] = a;
    red[NTH / 64 + (tid_ >> 6)] = b;
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

// A table of complex T (interleaved re, im) in global memory.
template <typename T> struct CTab {
  gptr<T> p;
  __device__ __forceinline__ V2<T> operator[](int i) const { return mk<T>(p[2 * i], p[2 * i + 1]); }
  __device__ __forceinline__ CTab operator+(int o) const { return CTab{p + 2 * o}; }
  __device__ __forceinline__ CTab operator-(int o) const { return CTab{p - 2 * o}; }
};
template <typename T>
__device__ __forceinline__ CTab<T> ctab(const void *p) { return CTab<T>{as_global<T>(p)}; }

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

// One pass of the Stockham DIF transform, in place in one LDS buffer: one
// stage of radix R1 (R2 == 1) or two consecutive stages, radix R1 at stride s
// then R2 at stride s R1, fused.  Butterfly b = p2 s + q loads the R1 R2
// inputs x[q + s(p2 + j2 sub2 + j sub1)], runs the R2 radix-R1 butterflies of
// the first stage (p = p2 + j2 sub2) with their twiddles w_cur^{pk}, then
// the R1 radix-R2 butterflies of the second (p2) with theirs, and stores
// y[q + s k + s R1 (R2 p2 + k2)]: the operations of fft/plan.nim:225-366 in
// the same order, without a round trip through LDS between the two stages.
//
// In place: every thread first loads the inputs of all its butterflies (at
// most MAXIT of them, in registers), the block waits, then every thread
// computes and stores; the caller waits again before the next pass.  One
// buffer per transform instead of two halves the LDS a row holds, so more
// rows share a CU.
//
// pad_in: x is the padded output of a first pass of s points (element i at
// i + i / s); pad_out (s == 1): y is written padded (one spare element after
// every R1 R2, so the lanes of a store hit different banks).  Both only
// change uniform strides and per-butterfly bases.
//
// LAST: the pass holds the transform's last stage.  Every stage but the last
// multiplies by twiddles, so the pass's twiddle multiplies are known at
// compile time (no branch around them: the compiler issues the twiddle
// loads of an iteration together instead of waiting on each group in turn).
template <typename T, int R1, int R2, int NTH, bool LAST, int BUDGET>
__device__ __forceinline__ void pass_inplace(V2<T> *__restrict__ x, CTab<T> tw1, CTab<T> tw2,
                                             int m, int s, int sub1, FastDiv sd, int pad_in,
                                             int pad_out) {
  constexpr int R = R1 * R2;
  constexpr int MAXIT = pass_iters_max(R, BUDGET);
  // the strides are uniform: keep them (and their multiples) in SGPRs
  s = uni(s);
  const int sub2 = uni(sub1 / R2);
  const int nb = uni(m / R);
  // sub1 > 1 and sub2 > 1 exactly when the stage is not the last
  constexpr bool twiddle1 = !LAST || R2 > 1;
  constexpr bool twiddle2 = R2 > 1 && !LAST;
  // input strides of j and j2 (in the padded layout when pad_in)
  const int in1 = uni(s * sub1 + pad_in * sub1), in2 = uni(s * sub2 + pad_in * sub2);
  const int out2 = uni(s * R1);  // output stride of k2
  const int in_p = uni(s + pad_in), out_p = uni(s * R + pad_out);
  const int t1_step = uni(sub2 * (R1 - 1));
  V2<T> v[MAXIT][R2][R1];
  {
    const int tid_ = opaque_tid();
#pragma unroll
    for (int it = 0; it < MAXIT; it++) {
      const int b = tid_ + NTH * it;
      if (b < nb) {
        const int p2 = (int)fdiv((uint32_t)b, sd);
        const int q = b - p2 * s;
        const V2<T> *xi = x + opaque_v(q + in_p * p2);
#pragma unroll
        for (int j2 = 0; j2 < R2; j2++)
#pragma unroll
          for (int j = 0; j < R1; j++) v[it][j2][j] = xi[j2 * in2 + j * in1];
      }
    }
  }
  __syncthreads();
  const int tid_ = opaque_tid();
#pragma unroll
  for (int it = 0; it < MAXIT; it++) {
    const int b = tid_ + NTH * it;
    if (b < nb) {
      const int p2 = (int)fdiv((uint32_t)b, sd);
      const int q = b - p2 * s;
      V2<T> *yo = x + opaque_v(q + out_p * p2);
      CTab<T> t1 = tw1 + p2 * (R1 - 1) - 1;
#pragma unroll
      for (int j2 = 0; j2 < R2; j2++) {
        Bfly<T, R1>::run(v[it][j2]);
        if constexpr (twiddle1) {
#pragma unroll
          for (int k = 1; k < R1; k++)
            v[it][j2][k] = cmul<T>(v[it][j2][k], t1[j2 * t1_step + k]);
        }
      }
      if constexpr (R2 == 1) {
#pragma unroll
        for (int k = 0; k < R1; k++) yo[k * s] = v[it][0][k];
      } else {
        CTab<T> t2 = tw2 + p2 * (R2 - 1) - 1;
#pragma unroll
        for (int k = 0; k < R1; k++) {
          V2<T> u[R2];
#pragma unroll
          for (int j2 = 0; j2 < R2; j2++) u[j2] = v[it][j2][k];
          Bfly<T, R2>::run(u);
          if constexpr (twiddle2) {
#pragma unroll
            for (int k2 = 1; k2 < R2; k2++) u[k2] = cmul<T>(u[k2], t2[k2]);
          }
#pragma unroll
          for (int k2 = 0; k2 < R2; k2++) yo[k * s + k2 * out2] = u[k2];
        }
      }
    }
    PSEG_ONE_AT_A_TIME();
  }
}

// Stage pairs whose butterfly holds at most FUSE points run fused (fusable,
// shared with the host); the registers of larger ones would cost more
// occupancy than the LDS round trip they save.  The host checks that no
// pass needs more than pass_iters_max butterflies per thread.
template <typename T, int R1, int NTH, int FUSE, bool LAST, int BUDGET>
__device__ __forceinline__ void pass_second(int r2, V2<T> *x, CTab<T> t1, CTab<T> t2, int m,
                                            int s, int sub1, FastDiv sd, int pi, int po) {
  switch (r2) {
    case 4: if constexpr (fusable(R1, 4, FUSE) && may_follow(R1, 4)) pass_inplace<T, R1, 4, NTH, LAST, BUDGET>(x, t1, t2, m, s, sub1, sd, pi, po); break;
    case 2: if constexpr (fusable(R1, 2, FUSE) && may_follow(R1, 2)) pass_inplace<T, R1, 2, NTH, LAST, BUDGET>(x, t1, t2, m, s, sub1, sd, pi, po); break;
    case 3: if constexpr (fusable(R1, 3, FUSE) && may_follow(R1, 3)) pass_inplace<T, R1, 3, NTH, LAST, BUDGET>(x, t1, t2, m, s, sub1, sd, pi, po); break;
    default: if constexpr (fusable(R1, 5, FUSE) && may_follow(R1, 5)) pass_inplace<T, R1, 5, NTH, LAST, BUDGET>(x, t1, t2, m, s, sub1, sd, pi, po); break;
  }
}

template <typename T, int NTH, int FUSE, bool LAST, int BUDGET>
__device__ __forceinline__ void pass_pair(int r, int r2, V2<T> *a, CTab<T> t, CTab<T> t2, int m,
                                          int s, int sub, FastDiv sd, int pi, int po) {
  switch (r) {
    case 4: pass_second<T, 4, NTH, FUSE, LAST, BUDGET>(r2, a, t, t2, m, s, sub, sd, pi, po); break;
    case 2: pass_second<T, 2, NTH, FUSE, LAST, BUDGET>(r2, a, t, t2, m, s, sub, sd, pi, po); break;
    case 3: pass_second<T, 3, NTH, FUSE, LAST, BUDGET>(r2, a, t, t2, m, s, sub, sd, pi, po); break;
    default: pass_second<T, 5, NTH, FUSE, LAST, BUDGET>(r2, a, t, t2, m, s, sub, sd, pi, po); break;
  }
}

template <typename T, int NTH, bool LAST, int BUDGET>
__device__ __forceinline__ void pass_single(int r, V2<T> *a, CTab<T> t, int m, int s, int sub,
                                            FastDiv sd, int pi, int po) {
  switch (r) {
    case 4: pass_inplace<T, 4, 1, NTH, LAST, BUDGET>(a, t, t, m, s, sub, sd, pi, po); break;
    case 2: pass_inplace<T, 2, 1, NTH, LAST, BUDGET>(a, t, t, m, s, sub, sd, pi, po); break;
    case 3: pass_inplace<T, 3, 1, NTH, LAST, BUDGET>(a, t, t, m, s, sub, sd, pi, po); break;
    default: pass_inplace<T, 5, 1, NTH, LAST, BUDGET>(a, t, t, m, s, sub, sd, pi, po); break;
  }
}

// ------------------------------------------- length-specialised passes
//
// For the transform lengths static_forward lists, the passes are compiled with
// the length's strides as constants: the LDS loads and stores take their
// offsets as immediates, the butterfly index divides by a constant, the
// iteration guards vanish where the butterflies fill every thread, and the
// twiddle offsets fold into the loads.  The butterflies, twiddles and their
// order are the generic pass's: the results are the same.

// radices_of (the host's plan): 4s, at most one 2, 3s, then 5s.
struct Radices {
  int n;
  int r[kMaxRadices];
};
__host__ __device__ constexpr Radices radices_const(int m) {
  Radices R{0, {}};
  while (m % 4 == 0) { R.r[R.n++] = 4; m /= 4; }
  if (m % 2 == 0) { R.r[R.n++] = 2; m /= 2; }
  while (m % 3 == 0) { R.r[R.n++] = 3; m /= 3; }
  while (m % 5 == 0) { R.r[R.n++] = 5; m /= 5; }
  return R;
}

// first_pass_pad of the host: the first pass's butterfly size when a second
// pass follows.
__host__ __device__ constexpr int first_pad_const(int m, int fuse) {
  const Radices R = radices_const(m);
  if (R.n < 2) return 0;
  const bool pair = fusable(R.r[0], R.r[1], fuse);
  const int used = pair ? 2 : 1;
  return R.n > used ? (pair ? R.r[0] * R.r[1] : R.r[0]) : 0;
}

// A twiddle of iteration `it`: from the registers loaded early (EARLY, the
// first iteration) or from the table.
template <bool EARLY, typename T, int N>
__device__ __forceinline__ V2<T> twiddle(int it, const V2<T> (&w)[N], int i, CTab<T> t, int j) {
  if constexpr (EARLY) {
    if (it == 0) return w[i];
  }
  return t[j];
}

template <typename T, int R1, int R2, int NTH, bool LAST, int M, int S, int SUB1, int PIN,
          int POUT, bool EARLY>
__device__ __forceinline__ void pass_static(V2<T> *__restrict__ x, CTab<T> tw1, CTab<T> tw2) {
  constexpr int R = R1 * R2;
  constexpr int SUB2 = SUB1 / R2;
  constexpr int NB = M / R;
  constexpr int IT = (NB + NTH - 1) / NTH;
  constexpr bool twiddle1 = !LAST || R2 > 1;
  constexpr bool twiddle2 = R2 > 1 && !LAST;
  constexpr int IN1 = S * SUB1 + PIN * SUB1, IN2 = S * SUB2 + PIN * SUB2;
  constexpr int OUT2 = S * R1;
  constexpr int INP = S + PIN, OUTP = S * R + POUT;
  constexpr int T1STEP = SUB2 * (R1 - 1);
  V2<T> v[IT][R2][R1];
  // EARLY: the first iteration's twiddles are loaded with the inputs, before
  // the block waits, so their latency (an L2 read) passes during the wait
  // instead of after it; the later iterations' as they run (registers).
  constexpr int NT1 = EARLY && twiddle1 ? R2 * (R1 - 1) : 0;
  constexpr int NT2 = EARLY && twiddle2 ? R2 - 1 : 0;
  V2<T> w1[NT1 > 0 ? NT1 : 1], w2[NT2 > 0 ? NT2 : 1];
  {
    const int tid_ = opaque_tid();
#pragma unroll
    for (int it = 0; it < IT; it++) {
      const int b = tid_ + NTH * it;
      if (NB % NTH == 0 || it + 1 < IT || b < NB) {
        const int p2 = b / S;
        const int q = b - p2 * S;
        const V2<T> *xi = x + (q + INP * p2);
#pragma unroll
        for (int j2 = 0; j2 < R2; j2++)
#pragma unroll
          for (int j = 0; j < R1; j++) v[it][j2][j] = xi[j2 * IN2 + j * IN1];
        if constexpr (EARLY) {
          if (it == 0) {
            CTab<T> t1 = tw1 + (p2 * (R1 - 1) - 1);
#pragma unroll
            for (int j2 = 0; j2 < R2; j2++)
#pragma unroll
              for (int k = 1; k < R1; k++)
                if constexpr (twiddle1) w1[j2 * (R1 - 1) + k - 1] = t1[j2 * T1STEP + k];
            CTab<T> t2 = tw2 + (p2 * (R2 - 1) - 1);
#pragma unroll
            for (int k2 = 1; k2 < R2; k2++)
              if constexpr (twiddle2) w2[k2 - 1] = t2[k2];
          }
        }
      }
    }
  }
  __syncthreads();
  const int tid_ = opaque_tid();
#pragma unroll
  for (int it = 0; it < IT; it++) {
    const int b = tid_ + NTH * it;
    if (NB % NTH == 0 || it + 1 < IT || b < NB) {
      const int p2 = b / S;
      const int q = b - p2 * S;
      V2<T> *yo = x + (q + OUTP * p2);
      CTab<T> t1 = tw1 + (p2 * (R1 - 1) - 1);
#pragma unroll
      for (int j2 = 0; j2 < R2; j2++) {
        Bfly<T, R1>::run(v[it][j2]);
        if constexpr (twiddle1) {
#pragma unroll
          for (int k = 1; k < R1; k++)
            v[it][j2][k] = cmul<T>(v[it][j2][k], twiddle<EARLY>(it, w1, j2 * (R1 - 1) + k - 1,
                                                                t1, j2 * T1STEP + k));
        }
      }
      if constexpr (R2 == 1) {
#pragma unroll
        for (int k = 0; k < R1; k++) yo[k * S] = v[it][0][k];
      } else {
        CTab<T> t2 = tw2 + (p2 * (R2 - 1) - 1);
#pragma unroll
        for (int k = 0; k < R1; k++) {
          V2<T> u[R2];
#pragma unroll
          for (int j2 = 0; j2 < R2; j2++) u[j2] = v[it][j2][k];
          Bfly<T, R2>::run(u);
          if constexpr (twiddle2) {
#pragma unroll
            for (int k2 = 1; k2 < R2; k2++)
              u[k2] = cmul<T>(u[k2], twiddle<EARLY>(it, w2, k2 - 1, t2, k2));
          }
#pragma unroll
          for (int k2 = 0; k2 < R2; k2++) yo[k * S + k2 * OUT2] = u[k2];
        }
      }
    }
    PSEG_ONE_AT_A_TIME();
  }
}

// The passes of the M-point transform from stage I (S points done).
template <typename T, int NTH, int FUSE, int M, int I, int S>
__device__ __forceinline__ void static_passes(const DevFft &f, V2<T> *a, CTab<T> tw) {
  constexpr Radices P = radices_const(M);
  if constexpr (I < P.n) {
    constexpr int r = P.r[I];
    constexpr int r2 = I + 1 < P.n ? P.r[I + 1] : 1;
    constexpr bool pair = I + 1 < P.n && fusable(r, r2, FUSE);
    constexpr int pad = first_pad_const(M, FUSE);
    constexpr int po = pad != 0 && I == 0, pi = pad != 0 && S != 1 && S == pad;
    constexpr int sub = M / (S * r);
    CTab<T> t = tw + uni(f.tw_off[I]);
    if constexpr (pair) {
      CTab<T> t2 = tw + uni(f.tw_off[I + 1]);
      pass_static<T, r, r2, NTH, I + 2 >= P.n, M, S, sub, pi, po, FUSE == kFusePipeline>(a, t, t2);
      __syncthreads();
      static_passes<T, NTH, FUSE, M, I + 2, S * r * r2>(f, a, tw);
    } else {
      pass_static<T, r, 1, NTH, I + 1 >= P.n, M, S, sub, pi, po, FUSE == kFusePipeline>(a, t, t);
      __syncthreads();
      static_passes<T, NTH, FUSE, M, I + 1, S * r>(f, a, tw);
    }
  }
}

// Whether every pass of the M-point transform fits NTH threads within the
// register budget (the host's fft_fits for the same plan).
template <int NTH, int FUSE, int BUDGET>
__host__ __device__ constexpr bool static_fits(int m) {
  const Radices P = radices_const(m);
  for (int i = 0; i < P.n;) {
    const bool pair = i + 1 < P.n && fusable(P.r[i], P.r[i + 1], FUSE);
    const int r = pair ? P.r[i] * P.r[i + 1] : P.r[i];
    if ((m / r + NTH - 1) / NTH > pass_iters_max(r, BUDGET)) return false;
    i += pair ? 2 : 1;
  }
  return true;
}

template <typename T, int NTH, int FUSE, int BUDGET, int M>
__device__ __forceinline__ bool try_static(const DevFft &f, V2<T> *a, int m) {
  if constexpr (static_fits<NTH, FUSE, BUDGET>(M)) {
    if (m == M) {
      static_passes<T, NTH, FUSE, M, 0, 1>(f, a, ctab<T>(uni_ptr(f.tw)));
      return true;
    }
  }
  return false;
}

// The complex lengths compiled with constant strides, in float32 kernels
// (each adds its passes to every such kernel): the row kernel's 1125, 1536
// and 1728 (real transforms of 2250, 3072 and 3456 points: band-passes of
// 2001- and 3001-sample traces at 2 ms) and the correlation's 8100 (16200
// points: 10001-sample records with a 6000-sample sweep).  Every other
// length runs the generic passes.
template <typename T, int NTH, int FUSE, int BUDGET>
__device__ __forceinline__ bool static_forward(const DevFft &f, V2<T> *a) {
  if constexpr (sizeof(T) == 4) {
    const int m = uni(f.m);
    if constexpr (FUSE == kFusePipeline && BUDGET == kBudgetPipeline) {
      if (try_static<T, NTH, FUSE, BUDGET, 1728>(f, a, m)) return true;
      if (try_static<T, NTH, FUSE, BUDGET, 1536>(f, a, m)) return true;
      if (try_static<T, NTH, FUSE, BUDGET, 1125>(f, a, m)) return true;
    } else {
      if (try_static<T, NTH, FUSE, BUDGET, 8100>(f, a, m)) return true;
    }
  }
  return false;
}

// The complex forward FFT of the m points at a, in place (a holds f.stride
// complex values): the plan's stages in passes, stage pairs of at most FUSE
// points fused (the host built f's padded layout with the same FUSE).
template <typename T, int NTH, int FUSE, int BUDGET = kBudgetCorr>
__device__ __forceinline__ void complex_forward(const DevFft &f, V2<T> *a) {
  if (static_forward<T, NTH, FUSE, BUDGET>(f, a)) return;
  CTab<T> tw = ctab<T>(uni_ptr(f.tw));
  int s = 1;
  const int nrad = uni(f.nrad), m = uni(f.m);
  const int padded = uni(f.pad) != 0;
  for (int i = 0; i < nrad;) {
    const int r = uni(f.rad[i]), sub = uni(f.sub[i]);
    const FastDiv sd = uni(f.sdiv[i]);
    CTab<T> t = tw + uni(f.tw_off[i]);
    const int r2 = i + 1 < nrad ? uni(f.rad[i + 1]) : 0;
    // the first pass writes the padded layout, the second reads it
    const int po = padded && i == 0, pi = padded && s != 1 && s == uni(f.pad);
    if (i + 1 < nrad && fusable(r, r2, FUSE)) {
      CTab<T> t2 = tw + uni(f.tw_off[i + 1]);
      if (i + 2 >= nrad) pass_pair<T, NTH, FUSE, true, BUDGET>(r, r2, a, t, t2, m, s, sub, sd, pi, po);
      else pass_pair<T, NTH, FUSE, false, BUDGET>(r, r2, a, t, t2, m, s, sub, sd, pi, po);
      s *= r * r2;
      i += 2;
    } else {
      if (i + 1 >= nrad) pass_single<T, NTH, true, BUDGET>(r, a, t, m, s, sub, sd, pi, po);
      else pass_single<T, NTH, false, BUDGET>(r, a, t, m, s, sub, sd, pi, po);
      s *= r;
      i += 1;
    }
    __syncthreads();
  }
}

// splitForward (fft/plan.nim:470-482), times 2: bin k of the real transform
// from Z[k] = a and Z[m-k] = b.  The factor 1/2 of the split, and the
// inverse's 1/n, are folded into the plan's bin multipliers (gain or pilot
// multiplier times 1/(2n), computed in float64 and rounded once).
template <typename T>
__device__ __forceinline__ V2<T> split_fwd2(V2<T> a, V2<T> b, V2<T> w) {
  T er = a.x + b.x;
  T ei = a.y - b.y;
  T orr = a.y + b.y;
  T oi = b.x - a.x;
  return mk<T>(er + w.x * orr - w.y * oi, ei + w.x * oi + w.y * orr);
}

// splitInverse (fft/plan.nim:484-499) with w = (cos, -sin): the swapped
// (Im Z', Re Z') of the inverse's input from X[k] = a and X[m-k] = b (the
// 1/n scale is in the multipliers).
template <typename T>
__device__ __forceinline__ V2<T> split_inv(V2<T> a, V2<T> b, V2<T> w) {
  T er = a.x + b.x, ei = a.y - b.y, dr = a.x - b.x, di = a.y + b.y;
  T wr = w.x, wi = -w.y;
  T orr = dr * wr - di * wi;
  T oi = dr * wi + di * wr;
  return mk<T>(ei + orr, er - oi);
}

// Bin multipliers of the fused split / multiply / unsplit pass.
template <typename T> struct GainMul {
  gptr<T> g;
  __device__ V2<T> operator()(V2<T> x, int k) const {
    T w = g[k];
    return mk<T>(x.x * w, x.y * w);
  }
};
template <typename T> struct CplxMul {
  CTab<T> mul;
  __device__ V2<T> operator()(V2<T> x, int k) const {
    V2<T> w = mul[k];
    return mk<T>(x.x * w.x - x.y * w.y, x.x * w.y + x.y * w.x);
  }
};

// Four T per bin pair in global memory (16- or 32-byte aligned rows).
template <typename T> struct Quad {
  T x, y, z, w;
};
template <typename T> struct QTab {
  gptr<T> p;
  __device__ __forceinline__ Quad<T> operator[](int i) const {
    gptr<T> q = (gptr<T>)__builtin_assume_aligned((const T *)(p + 4 * i), 4 * sizeof(T));
    return Quad<T>{q[0], q[1], q[2], q[3]};
  }
};

// The split, a per-bin multiplier and the unsplit of the inverse, folded into
// one linear map per bin pair (k, m - k): the three steps of `combine` below
// with GainMul or CplxMul are, in exact arithmetic,
//   W[k] = swap(alpha_k Z[k] + beta_k conj(Z[m - k]))
// with w = exp(-2 pi i k / n), s = Im w, c = Re w and the multipliers M:
//   alpha_k = M_k (2 + 2s) + conj(M_{m-k}) (2 - 2s),
//   beta_k  = 2 i c (M_k - conj(M_{m-k})).
// For a real gain alpha is real and beta imaginary, and beta_{m-k} =
// beta_k, so a bin pair costs 8 multiply-adds instead of about 40 (the
// split's and unsplit's twiddles and the multiplies).  The host computes the
// coefficients in float64 and rounds them once; bin 0 and the Nyquist bin
// keep the split's form (both real).
//
// Real gains: coef[0] = (M_0, M_m, 0, 0); coef[k] = (alpha_k, alpha_{m-k},
// Im beta_k, 0) for k = 1 .. m/2.
template <typename T, int NTH>
__device__ __forceinline__ void combine_gain(const DevFft &f, V2<T> *A, QTab<T> coef) {
  const int tid_ = opaque_tid();
  const int m = uni(f.m);
  for (int k = tid_; k <= m / 2; k += NTH) {
    const Quad<T> c = coef[k];
    if (k == 0) {
      const V2<T> z = A[0];
      const T x0 = (z.x + z.y) * c.x, xm = (z.x - z.y) * c.y;
      A[0] = mk<T>((x0 - xm) * (T)2, (x0 + xm) * (T)2);
    } else {
      const int kk = m - k;
      const V2<T> a = A[k], b = A[kk];
      // alpha a + i beta conj(b), swapped
      A[k] = mk<T>(c.x * a.y + c.z * b.x, c.x * a.x + c.z * b.y);
      if (kk != k) A[kk] = mk<T>(c.y * b.y + c.z * a.x, c.y * b.x + c.z * a.y);
    }
  }
}

// Complex multipliers: coef[0] = (Re M_0, Re M_m, 0, 0), coef[1] unused;
// coef[2k] = (alpha_k, beta_k), coef[2k + 1] = (alpha_{m-k}, beta_{m-k}) as
// (re, im, re, im) for k = 1 .. m/2.
template <typename T, int NTH>
__device__ __forceinline__ void combine_cplx(const DevFft &f, V2<T> *A, QTab<T> coef) {
  const int tid_ = opaque_tid();
  const int m = uni(f.m);
  for (int k = tid_; k <= m / 2; k += NTH) {
    if (k == 0) {
      const Quad<T> c = coef[0];
      const V2<T> z = A[0];
      const T x0 = (z.x + z.y) * c.x, xm = (z.x - z.y) * c.y;
      A[0] = mk<T>((x0 - xm) * (T)2, (x0 + xm) * (T)2);
    } else {
      const int kk = m - k;
      const V2<T> a = A[k], b = A[kk];
      const Quad<T> c = coef[2 * k];
      // alpha a + beta conj(b), swapped
      A[k] = mk<T>(c.x * a.y + c.y * a.x + c.w * b.x - c.z * b.y,
                   c.x * a.x - c.y * a.y + c.z * b.x + c.w * b.y);
      if (kk != k) {
        const Quad<T> e = coef[2 * k + 1];
        A[kk] = mk<T>(e.x * b.y + e.y * b.x + e.w * a.x - e.z * a.y,
                      e.x * b.x - e.y * b.y + e.z * a.x + e.w * a.y);
      }
    }
  }
}

// Z (the forward complex FFT of the packed samples) -> W, the swapped input
// of the inverse: split into bins, multiply, unsplit.  W may be Z: each
// thread reads bins k and m - k before it writes them.  The split table is
// read once per bin pair: exp(-2 pi i (m - k) / n) = -conj(exp(-2 pi i k / n))
// for n = 2m.
template <typename T, int NTH, typename Mul>
__device__ __forceinline__ void combine(const DevFft &f, const V2<T> *Z, V2<T> *W, Mul mul) {
  const int tid_ = opaque_tid();
  const int m = uni(f.m);
  CTab<T> sp = ctab<T>(uni_ptr(f.split));
  for (int k = tid_; k <= m / 2; k += NTH) {
    if (k == 0) {
      V2<T> z = Z[0];
      V2<T> x0 = mul(mk<T>(z.x + z.y, (T)0), 0);
      V2<T> xm = mul(mk<T>(z.x - z.y, (T)0), m);
      W[0] = mk<T>((x0.x - xm.x) * (T)2, (x0.x + xm.x) * (T)2);
    } else {
      const int kk = m - k;
      V2<T> a = Z[k], b = Z[kk];
      const V2<T> w = sp[k];
      V2<T> xk = mul(split_fwd2<T>(a, b, w), k);
      if (kk != k) {
        const V2<T> wk = mk<T>(-w.x, w.y);
        V2<T> xkk = mul(split_fwd2<T>(b, a, wk), kk);
        W[k] = split_inv<T>(xk, xkk, w);
        W[kk] = split_inv<T>(xkk, xk, wk);
      } else {
        W[k] = split_inv<T>(xk, xk, w);
      }
    }
  }
}

// ------------------------------------------------------------ row helpers

template <typename T, int PER>
__device__ __forceinline__ double row_sum(const T (&v)[PER]) {
  double s = 0;  // slots past the live end hold +0: adding them changes nothing
#pragma unroll
  for (int k = 0; k < PER; k++) s += (double)v[k];
  return s;
}

// Least-squares line (filter/edges.nim:16-30, core/basic.nim:61-81 share it).
template <typename T, int NTH, int PER>
__device__ __forceinline__ void trend_of(const T (&v)[PER], int n, int trend, RedSlots &red, double &a,
                         double &b) {
  const int tid_ = opaque_tid();
  if (trend == 0) {
    a = 0;
    b = 0;
    return;
  }
  double sx = 0, skx = 0;  // slots past the live end hold +0
#pragma unroll
  for (int k = 0; k < PER; k++) {
    sx += (double)v[k];
    skx += (double)k * (double)v[k];
  }
  if (trend == 1) {
    a = block_sum<NTH>(sx, red.next()) / (double)n;
    b = 0;
    return;
  }
  // sum of j x_j over this thread's slots j = tid + NTH k
  double sxi = (double)tid_ * sx + (double)NTH * skx;
  block_sum2<NTH>(sx, sxi, red.next());
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
template <typename T, int NTH, int PER>
__device__ __forceinline__ void filter_step(const DevFilter &F, T (&v)[PER], int n, char *lds, RedSlots &red) {
  const int tid_ = opaque_tid();
  const DevFft &f = F.fft;
  T *buf = (T *)lds;                          // nfft samples = m complex
  V2<T> *A = (V2<T> *)lds;                   // the transform's buffer (f.stride)
  const int fn = uni(f.n);
  double a, b;
  trend_of<T, NTH, PER>(v, n, uni(F.trend), red, a, b);
  const int p = uni(F.pad), nf = fn;
  const int reflect = uni(F.reflect);
  // pack (packInto, filter/apply.nim:194-216): detrended data ...  The line
  // a + b j at slot k is (a + b tid) + (b NTH) k.
  {
    const double a0 = a + b * (double)tid_, bs = b * (double)NTH;
    T *bp = buf + p + tid_;
#pragma unroll
    for (int k = 0; k < PER; k++) {
      int j = tid_ + NTH * k;
      if (j < n) bp[NTH * k] = (T)((double)v[k] - (a0 + bs * (double)k));
      PSEG_ONE_AT_A_TIME();
    }
  }
  if (!reflect) {
    for (int j = tid_; j < p; j += NTH) buf[j] = (T)0;
    for (int j = p + n + tid_; j < nf; j += NTH) buf[j] = (T)0;
  } else {
    for (int j = 2 * p + n + tid_; j < nf; j += NTH) buf[j] = (T)0;
  }
  __syncthreads();
  if (reflect) {
    // ... and odd reflections 2 x_edge - x_k tapered by fall[k - 1]
    gptr<T> fall = as_global<T>(uni_ptr(F.fall));
    const T d0 = buf[p], dn = buf[p + n - 1];
    for (int k = 1 + tid_; k <= p; k += NTH) {
      int mk_ = k < n - 1 ? k : n - 1;
      T fk = fall[k - 1];
      T rl = buf[p + mk_];
      T rr = buf[p + n - 1 - mk_];
      buf[p - k] = (d0 + d0 - rl) * fk;
      buf[p + n - 1 + k] = (dn + dn - rr) * fk;
    }
    __syncthreads();
  }
  // forward FFT, split * gain * unsplit, inverse FFT, all in place: one
  // copy of the FFT code serves both directions
#pragma nounroll
  for (int dir = 0; dir < 2; dir++) {
    complex_forward<T, NTH, kFusePipeline, kBudgetPipeline>(f, A);
    if (dir == 0) {
      combine_gain<T, NTH>(f, A, QTab<T>{as_global<T>(uni_ptr(F.coef))});
      __syncthreads();
    }
  }
  const T *out = (const T *)A;
  // unpack (unpackFrom, filter/apply.nim:218-231): sample i of the inverse is
  // element i ^ 1 of the swapped result.
  const double dcg = uni(F.dc_gain);
  double gb = b * dcg;
  double ga = uni(F.zero_dc) ? -gb * (double)(n - 1) / 2 : a * dcg;
  const bool plain = ga == 0 && gb == 0;
  // Every slot is written (+0 past the live end: those samples were stored
  // unchanged before the steps), so the row's registers are dead during the
  // transforms.  The indices come from a fresh thread id: reusing the pack's
  // would keep them live across both transforms.
  const int tid2 = opaque_tid();
  // Element (p + j) ^ 1 for j = tid + NTH k is ((p + tid) ^ 1) + NTH k (NTH
  // is even): one base, immediate offsets, and every read of the row in
  // flight at once.  Slots past the live end read whatever lies there (a
  // filter step's LDS covers p plus the widest row of NTH threads, plus 2).
  {
    const T *ob = out + ((p + tid2) ^ 1);
#pragma unroll
    for (int k = 0; k < PER; k++) v[k] = ob[NTH * k];
  }
  if (plain) {  // no trend to add back (a 0 Hz gain of 0, or no trend)
#pragma unroll
    for (int k = 0; k < PER; k++) {
      int j = tid2 + NTH * k;
      v[k] = j < n ? v[k] : (T)0;
    }
  } else {
    const double g0 = ga + gb * (double)tid2, gs = gb * (double)NTH;
#pragma unroll
    for (int k = 0; k < PER; k++) {
      int j = tid2 + NTH * k;
      T y = (T)((double)v[k] + (g0 + gs * (double)k));
      v[k] = j < n ? y : (T)0;
      PSEG_ONE_AT_A_TIME();
    }
  }
  __syncthreads();
}

// A segmented scan value: a float64 sum and a flag.  Forward (prefix)
// combine(L, R) = (R.f ? R.v : L.v + R.v, L.f | R.f): R's flag means a block
// starts in R, so nothing left of it carries over.  Backward (suffix)
// combine(L, R) = (L.v + (L.f ? 0 : R.v), L.f | R.f).
struct Seg {
  double v;
  int f;
};

// Exclusive segmented scan over the NTH threads of the block in thread order
// (dir 0: from thread 0 up, prefix; dir 1: from the last thread down,
// suffix).  red: 2 NTH / 64 doubles of LDS.
template <int NTH>
__device__ __forceinline__ Seg block_seg_scan_excl(Seg x, int dir, double *red) {
  const int tid_ = opaque_tid();
  const int lane = tid_ & 63, wave = tid_ >> 6;
  Seg incl = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    double ov = dir == 0 ? __shfl_up(incl.v, d, 64) : __shfl_down(incl.v, d, 64);
    int of = dir == 0 ? __shfl_up(incl.f, d, 64) : __shfl_down(incl.f, d, 64);
    bool take = dir == 0 ? lane >= d : lane + d < 64;
    if (take) {
      if (!incl.f) incl.v = ov + incl.v;
      incl.f |= of;
    }
  }
  // this lane's exclusive value within the wave
  Seg ex;
  ex.v = dir == 0 ? __shfl_up(incl.v, 1, 64) : __shfl_down(incl.v, 1, 64);
  ex.f = dir == 0 ? __shfl_up(incl.f, 1, 64) : __shfl_down(incl.f, 1, 64);
  const bool first = dir == 0 ? lane == 0 : lane == 63;
  if (first) {
    ex.v = 0;
    ex.f = 0;
  }
  // the wave aggregates, then the waves before this one (in scan order)
  if ((dir == 0 && lane == 63) || (dir == 1 && lane == 0)) {
    red[wave] = incl.v;
    red[(NTH / 64) + wave] = (double)incl.f;
  }
  __syncthreads();
  Seg carry;
  carry.v = 0;
  carry.f = 0;
  if (dir == 0) {
    for (int w = 0; w < wave; w++) {
      int wf = red[(NTH / 64) + w] != 0;
      carry.v = wf ? red[w] : carry.v + red[w];
      carry.f |= wf;
    }
    // combine(carry, ex)
    Seg r;
    r.v = ex.f ? ex.v : carry.v + ex.v;
    r.f = carry.f | ex.f;
    return r;
  }
  for (int w = (NTH / 64) - 1; w > wave; w--) {
    int wf = red[(NTH / 64) + w] != 0;
    carry.v = red[w] + (wf ? 0.0 : carry.v);
    carry.f |= wf;
  }
  // combine(ex, carry)
  Seg r;
  r.v = ex.v + (ex.f ? 0.0 : carry.v);
  r.f = ex.f | carry.f;
  return r;
}

// The AGC gain step's level and quotient.  float32: the hardware square
// root and reciprocal (within 1 ulp each; the gain then differs from the
// correctly rounded one by at most a few ulp, far inside the float32
// tolerance) instead of the correctly rounded sequences, which cost about
// ten instructions each per sample.  float64: correctly rounded.
__device__ __forceinline__ float agc_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double agc_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float agc_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double agc_div(double a, double b) { return a / b; }

// AGC (gain/gain.nim:404-461) on the row in registers.  The window sums are
// the reference's block prefix and suffix sums (blocks of w = 2h + 1 samples,
// gain/gain.nim:351-390) computed by the whole block: each thread takes a
// contiguous run of samples, segmented scans carry the float64 sums across
// runs, and no sum ever subtracts.
template <typename T, int NTH, int PER>
__device__ __forceinline__ void agc_step(const DevAgc &G, T (&v)[PER], int n, char *lds,
                                         RedSlots &red) {
  const int tid_ = opaque_tid();
  const int mode = uni(G.mode);
  const double gtarget = uni(G.target), geps = uni(G.epsilon);
  double sq = 0;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    sq += (double)v[k] * (double)v[k];  // +0 past the live end
  }
  const double rms = sqrt(block_sum<NTH>(sq, red.next()) / (double)n);
  if (rms == 0) return;  // all zero: nothing to scale
  if (!fits_in<T>(gtarget / rms / geps)) return;
  const int h = uni(G.half), w = 2 * h + 1;
  // The reference's five ranges of i (gain/gain.nim:404-461).
  const int first = min(h, n);
  const int last = max(first, n - h);
  const int headCut = min(max(n - h, 0), first);
  const int lastBlock = (n - 1) / w * w;
  const int tailCut = min(max(lastBlock + h, last), n);
  // The window sums run on x / rms (so x^2 * unit, unit = 1 / rms^2 in rms
  // mode, |x| * unit, unit = 1 / rms, in mean |x| mode): every term is at
  // most n, so the sums stored as T neither overflow nor fall into
  // subnormals whatever the trace's amplitude, and the means are the
  // reference's scaled means (gain/gain.nim:430-461) with no further scale.
  const double ir = 1 / rms;
  const T invW = (T)(1.0 / (double)w);
  // LDS, n + first + 1 values: s[0, n) holds x, then the suffix sums S in
  // place, then over s[0, n - w] the interior means (S[e] + P[e + w]) / w in
  // T; hp[0, headCut) holds the prefix sums P[h + 1 + i] the head range
  // reads, hp[first] holds P[n].  The tail reads S[e] at e > n - w, which the
  // means never overwrite.
  T *s = (T *)lds;
  T *hp = s + n;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    int j = tid_ + NTH * k;
    if (j < n) s[j] = v[k];
  }
  __syncthreads();
  // Runs of c contiguous samples, c odd: thread t's run starts at t·c, so the
  // lanes of a wave touch LDS banks t·c mod 32, all different (an even c
  // such as 24 puts 8 lanes on each bank); an even ceil(n / NTH) rounds up
  // by one, so c <= PER + 1.
  const int c = ((n + NTH - 1) / NTH) | 1;
  const int c0 = min((int)tid_ * c, n), c1 = min(c0 + c, n);
  constexpr int RUN = PER + 1;
  double tv[RUN];
  Seg pre, suf;
  pre.v = 0;
  pre.f = 0;
  suf.v = 0;
  suf.f = 0;
  // e % w, kept incrementally (no division per sample)
  const int r0 = c0 < n ? c0 % w : 0;
  int r = r0;
  if (c <= w) {
    // A run holds at most one block start, at ks: the suffix part is the
    // terms before it, the prefix part the terms from it (the same sums, in
    // the same order, as the general loop below).  The run's samples are
    // read at once (one wait), and the mode picks one of two loops.
    const int len = c1 - c0;
    const int ks = r0 == 0 ? 0 : w - r0;
    double lo = 0, hi = 0;
    auto terms = [&](auto rms_mode) {
#pragma unroll
      for (int k = 0; k < RUN; k++) {
        const double x = (double)(c0 + k < c1 ? s[c0 + k] : (T)0) * ir;  // 0 past the run
        const double t = decltype(rms_mode)::value ? x * x : fabs(x);
        tv[k] = t;
        if (k < ks) lo += t;
        else hi += t;
        PSEG_ONE_AT_A_TIME();
      }
    };
    if (mode == 0) terms(std::true_type{});
    else terms(std::false_type{});
    const bool has = ks < len;
    pre.v = has ? hi : lo;
    pre.f = has;
    suf.v = lo;
    suf.f = has;
  } else {
    bool past_start = false;  // a block start seen after c0 (suffix)
#pragma unroll
    for (int k = 0; k < RUN; k++) {
      const int e = c0 + k;
      double t = 0;
      if (e < c1) {
        const double x = (double)s[e] * ir;
        t = mode == 0 ? x * x : fabs(x);
        const bool start = r == 0;
        // prefix: the sum since the last block start in the run
        pre.v = start ? t : pre.v + t;
        pre.f |= start;
        // suffix: the run's terms before the first block start after c0
        if (start && k > 0) past_start = true;
        if (!past_start) suf.v += t;
        r = r + 1 == w ? 0 : r + 1;
      }
      tv[k] = t;
      PSEG_ONE_AT_A_TIME();
    }
    suf.f = (c0 < c1) && (r0 == 0 || past_start);
    if (c0 < c1 && r0 == 0) suf.v = 0;
  }
  Seg cpre = block_seg_scan_excl<NTH>(pre, 0, red.next());
  Seg csuf = block_seg_scan_excl<NTH>(suf, 1, red.next());
  // suffix sums S[e], over x in place (this thread's run only); the run's
  // last sample has remainder (c1 - 1) % w = (r0 + len - 1) % w
  const int len = c1 - c0;
  double run = (c1 < n && (r0 + len) % w != 0) ? csuf.v : 0.0;
  // A run of c <= w samples holds at most one block start, at k = ks, and
  // its prefix sums are 0 (a block end) only at k + 1 = zk; longer runs
  // (windows under c samples) track e % w sample by sample.
  const int ks = r0 == 0 ? 0 : w - r0;
  const int zk = ks == 0 ? w : ks;
  if (c <= w) {
#pragma unroll
    for (int k = RUN - 1; k >= 0; k--) {
      if (k < len) {
        run += tv[k];
        s[c0 + k] = (T)run;
        if (k == ks) run = 0;
      }
    }
  } else {
    r = len > 0 ? (r0 + len - 1) % w : 0;
#pragma unroll
    for (int k = RUN - 1; k >= 0; k--) {
      if (k < len) {
        run += tv[k];
        s[c0 + k] = (T)run;
        if (r == 0) run = 0;
        r = r == 0 ? w - 1 : r - 1;
      }
    }
  }
  __syncthreads();
  // prefix sums P[j], j = e + 1 (0 at a block end), used as they are made:
  // the interior mean of sample j - h - 1 replaces S[j - w], which only this
  // thread reads; the head's go to hp, and P[n] (the last run's last) to
  // hp[first].
  run = cpre.v;
  r = r0;
  T plast = (T)0;
  const int jh = c0 - h;  // j - h - 1 at k = 0
#pragma unroll
  for (int k = 0; k < RUN; k++) {
    if (k < len) {
      T pj;
      if (c <= w) {
        if (k == ks) run = 0;
        run += tv[k];
        pj = k + 1 == zk ? (T)0 : (T)run;
      } else {
        if (r == 0) run = 0;
        run += tv[k];
        r = r + 1 == w ? 0 : r + 1;
        pj = r == 0 ? (T)0 : (T)run;
      }
      const int q = c0 + k + 1 - w;  // j - w
      if (q >= 0) s[q] = (s[q] + pj) * invW;
      if ((unsigned)(jh + k) < (unsigned)headCut) hp[jh + k] = pj;
      plast = pj;
    }
  }
  if (len > 0 && c1 == n) hp[first] = plast;
  __syncthreads();
  // The gain step in T on means scaled by the trace level.
  const T kk = (T)(gtarget / rms);
  const T eps = (T)geps;
  const double pn = (double)hp[first];
  // 1 / count for a window of `count` samples: inv[c] = 1/c; the tail's
  // rev[m] = 1/(2h+1-m) is inv[2h+1-m], the same float64 value.
  gptr<double> inv = as_global<double>(uni_ptr(G.inv));
  const double headWhole = headCut < first ? pn * inv[n] : 0.0;
  const int tid3 = opaque_tid();  // fresh: nothing per slot stays live across the scans
  const int wave0 = __builtin_amdgcn_readfirstlane(tid3);  // the wave's first thread
  // s[i - h] of every slot, read at once: the interior mean, or the suffix
  // sum an edge range adds (clamped into the row for slots that use neither)
  T sm[PER];
#pragma unroll
  for (int k = 0; k < PER; k++) sm[k] = s[min(max(tid3 + NTH * k - h, 0), n - 1)];
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int i = tid3 + NTH * k;
    const int lo = wave0 + NTH * k;  // the wave's samples: lo .. lo + 63
    if (lo >= first && lo + 63 < last) {
      // the whole wave in the interior range (most of the row): full windows
      const T mean = sm[k];
      const T lvl = mode == 0 ? agc_sqrt(mean) : mean;
      const T gi = agc_div(kk, lvl + eps);
      v[k] = (T)(v[k] * gi);
    } else if (i < n) {
      // the interior mean as stored; the edge ranges as one expression,
      // mean = (sv + pv) * M with sv, pv 0 where a range does not add them
      // (x + 0 = x exactly)
      const bool head = i < headCut, mid = i >= first && i < last;
      const bool whole = i >= headCut && i < first, tail1 = i >= last && i < tailCut;
      T mean;
      if (mid) {
        mean = sm[k];
      } else {
        const double sv = i >= first ? (double)sm[k] : 0.0;
        const double pv = head ? (double)hp[i] : (tail1 ? pn : 0.0);
        const int c = head ? i + h + 1 : n - i + h;
        const double M = inv[min(max(c, 1), w)];
        mean = (T)(whole ? headWhole : (sv + pv) * M);
      }
      const T lvl = mode == 0 ? agc_sqrt(mean) : mean;
      const T gi = agc_div(kk, lvl + eps);
      v[k] = (T)(v[k] * gi);
    }
    PSEG_ONE_AT_A_TIME();
  }
  __syncthreads();
}

// Exclusive sum of x over the block's threads in thread order (wave
// shuffles, then the waves' totals through LDS).  wsum: NTH / 64 values.
template <int NTH>
__device__ __forceinline__ uint32_t block_excl_scan_u32(uint32_t x, uint32_t *wsum) {
  const int tid_ = opaque_tid();
  const int lane = tid_ & 63, wave = tid_ >> 6;
  uint32_t incl = x;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const uint32_t o = __shfl_up(incl, d, 64);
    if (lane >= d) incl += o;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  uint32_t before = 0;
  for (int w = 0; w < wave; w++) before += wsum[w];
  return before + incl - x;
}

// The k-th smallest of the row's |x| bit patterns (exact), by radix select
// over 8-bit digits (the last one 7 bits) with an LDS histogram.
template <typename T, int NTH, int PER>
__device__ __forceinline__ typename Bits<T>::S select_bits(const T (&v)[PER], int n, int64_t rank,
                                           uint32_t *hist, int64_t *shared_i) {
  const int tid_ = opaque_tid();
  using S = typename Bits<T>::S;
  using U = typename Bits<T>::U;
  __shared__ uint32_t wsum[NTH / 64];
  U prefix = 0, mask = 0;
  int64_t r = rank;
  // digits from just below the cleared sign bit: the first is the exponent,
  // so each bin holds one binade (fewer colliding atomics than two)
  for (int top = Bits<T>::width - 1; top > 0;) {
    const int shift = max(top - 8, 0);
    const U dmask = ((U)1 << (top - shift)) - 1;
    top = shift;
    for (int i = tid_; i < 256; i += NTH) hist[i] = 0;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; k++) {
      int j = tid_ + NTH * k;
      if (j < n) {
        U key = (U)Bits<T>::abs_bits(v[k]);
        if ((key & mask) == prefix) atomicAdd(&hist[(uint32_t)((key >> shift) & dmask)], 1u);
      }
    }
    __syncthreads();
    // The digit holding rank r: an exclusive scan of the 256 counts over the
    // block (thread t holds bins t BPT .. t BPT + BPT - 1), and the one bin
    // with below <= r < below + count reports it.
    {
      constexpr int BPT = NTH >= 256 ? 1 : 256 / NTH;
      uint32_t c[BPT];
      uint32_t mine = 0;
#pragma unroll
      for (int i = 0; i < BPT; i++) {
        const int bin = tid_ * BPT + i;
        c[i] = bin < 256 ? hist[bin] : 0u;
        mine += c[i];
      }
      uint32_t below = block_excl_scan_u32<NTH>(mine, wsum);
      const uint32_t rr = (uint32_t)r;
#pragma unroll
      for (int i = 0; i < BPT; i++) {
        if (below <= rr && rr < below + c[i]) {
          shared_i[0] = tid_ * BPT + i;
          shared_i[1] = (int64_t)(rr - below);
        }
        below += c[i];
      }
    }
    __syncthreads();
    prefix |= (U)shared_i[0] << shift;
    mask |= dmask << shift;
    r = shared_i[1];
    __syncthreads();
  }
  return (S)prefix;
}

// muteTimeMs (core/mute.nim:100-118) of a trace at `offset`.
__device__ __forceinline__ double mute_time_ms(const DevMute &M, double offset) {
  const int kind = uni(M.kind);
  if (kind == 0) return uni(M.t_ms);
  if (kind == 2) return uni(M.t_ms) + fabs(offset) * uni(M.ms_per_dist) + uni(M.pad_ms);
  const double x = fabs(offset);
  if (isnan(x)) return x;
  gptr<double> o = as_global<double>(uni_ptr(M.offsets));
  gptr<double> t = as_global<double>(uni_ptr(M.times));
  const int last = uni(M.knots) - 1;
  if (x <= o[0]) return t[0];
  if (x >= o[last]) return t[last];
  int lo = 0, hi = last;  // the last knot <= x, by bisection
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (o[mid] <= x) lo = mid;
    else hi = mid;
  }
  const double u = (x - o[lo]) / (o[hi] - o[lo]);
  return t[lo] + u * (t[hi] - t[lo]);
}

// firstAtOrAfter / firstAfter (core/mute.nim:122-132).
__device__ __forceinline__ int first_at_or_after(double c, int n) {
  return !(c > 0) ? 0 : c >= (double)n ? n : (int)ceil(c);
}
__device__ __forceinline__ int first_after(double c, int n) {
  return !(c >= 0) ? 0 : c >= (double)n ? n : (int)floor(c) + 1;
}

// sin x for 0 <= x <= π/2 (a mute taper's argument): the Taylor series to
// x^21, whose remainder there is below 2e-18 (under half an ulp of the
// result), by Horner's rule in x² — a few FMAs and registers where the
// general sin's range reduction would cost the row kernel its registers.
// A float64 constant made in SGPRs where it is used (an asm barrier): the
// compiler would otherwise hoist the coefficients below out of the step loop
// into twenty VGPRs.
__device__ __forceinline__ double sconst(double c) {
  asm volatile("" : "+s"(c));
  return c;
}
__device__ __forceinline__ double taper_sin(double x) {
  const double z = x * x;
  double p = sconst(1.0 / 51090942171709440000.0);           // 1/21!
  p = fma(p, z, sconst(-1.0 / 121645100408832000.0));        // 1/19!
  p = fma(p, z, sconst(1.0 / 355687428096000.0));            // 1/17!
  p = fma(p, z, sconst(-1.0 / 1307674368000.0));             // 1/15!
  p = fma(p, z, sconst(1.0 / 6227020800.0));                 // 1/13!
  p = fma(p, z, sconst(-1.0 / 39916800.0));                  // 1/11!
  p = fma(p, z, sconst(1.0 / 362880.0));                     // 1/9!
  p = fma(p, z, sconst(-1.0 / 5040.0));                      // 1/7!
  p = fma(p, z, sconst(1.0 / 120.0));                        // 1/5!
  p = fma(p, z, sconst(-1.0 / 6.0));                         // 1/3!
  return fma(x * z, p, x);
}

// muteTop / muteBottom (core/mute.nim:136-180) of the row's live part at
// column `col` (the row's column of the mute time, muteCol, core/mute.nim:
// 189-194), the taper in columns, the weights sin² in float64 (taper_sin
// and the argument as (π/2 / taper) x: within a few ulp of the reference's
// sin(π/2 x / taper), far inside the float32 rounding of the product).
template <typename T, int NTH, int PER>
__device__ __forceinline__ void mute_step(const DevMute &M, bool top, T (&v)[PER], int n,
                                          double col) {
  if (isnan(col)) return;  // a NaN time or t0Ms: the row is left alone
  const int tid_ = opaque_tid();
  const double dt = uni(M.dt_ms);
  const double tc = uni(M.taper_ms) / dt;
  const double taper = tc < INFINITY ? tc : 0.0;
  constexpr double kHalfPi = 1.5707963267948966;  // π / 2 as the float64 nearest
  const double arg_per_col = taper > 0 ? kHalfPi / taper : 0.0;  // one division per row
  if (top) {
    const double ramp_start = col - taper;
    const int zero_end = first_at_or_after(ramp_start, n);
    const int e = taper > 0 ? first_at_or_after(col, n) : zero_end;
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int j = tid_ + NTH * k;
      if (j < zero_end) {
        v[k] = (T)0;
      } else if (j < e) {
        const double s = taper_sin(arg_per_col * ((double)j - ramp_start));
        v[k] = (T)((double)v[k] * (s * s));
      }
      PSEG_ONE_AT_A_TIME();
    }
  } else {
    const double ramp_end = col + taper;
    const int zero_start = first_after(ramp_end, n);
    const int b = taper > 0 ? first_after(col, n) : zero_start;
#pragma unroll
    for (int k = 0; k < PER; k++) {
      const int j = tid_ + NTH * k;
      if (j < n) {
        if (j >= zero_start) {
          v[k] = (T)0;
        } else if (j >= b) {
          const double s = taper_sin(arg_per_col * (ramp_end - (double)j));
          v[k] = (T)((double)v[k] * (s * s));
        }
      }
      PSEG_ONE_AT_A_TIME();
    }
  }
}

// pow as a call of its own: inlined into the row kernel, its polynomial
// constants are hoisted out of the loop over the steps into VGPRs and spill.
__device__ __attribute__((noinline)) double pow_call(double b, double e) { return pow(b, e); }

// gainDbAt / factorAt (gain/curves.nim:134-160) in float64, without
// contraction, as the curve's own arithmetic: linear in dB between knots
// (bisection) and flat outside them, 10^(dB/20); (max(t, minMs)/refMs)^power
// directly for a power curve.
__device__ __forceinline__ double gain_factor_at(const DevGain &G, double t) {
#pragma clang fp contract(off)
  if (isnan(t)) return t;
  const int kind = uni(G.kind);
  if (kind == 1) return pow_call(fmax(t, uni(G.min_ms)) / uni(G.ref_ms), uni(G.power));
  double db;
  if (kind == 2) {
    db = uni(G.db_per_s) * t / 1000;
  } else {
    gptr<double> kt = as_global<double>(uni_ptr(G.knot_ms));
    gptr<double> kd = as_global<double>(uni_ptr(G.knot_db));
    const int last = uni(G.knots) - 1;
    if (t <= kt[0]) {
      db = kd[0];
    } else if (t >= kt[last]) {
      db = kd[last];
    } else {
      int lo = 0, hi = last;  // kt[lo] < t < kt[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (kt[mid] <= t) lo = mid;
        else hi = mid;
      }
      const double f = (t - kt[lo]) / (kt[hi] - kt[lo]);
      db = kd[lo] + f * (kd[hi] - kd[lo]);
    }
  }
  return pow_call(10.0, db / 20);
}

// applyGainCurve's workspace form on the row's live part (gain/curves.nim:
// 340-355, 380-416), sample 0 at `start` ms: where the row lies on the table
// (spanOf, computed once per row in SGPRs), its entries (two interpolated in T
// off the grid); elsewhere T(factorAt(t)), a zero staying zero; a NaN start
// leaves the row alone.  The row goes through LDS: each lane stores its
// samples, scales them there in a loop that is not unrolled (one copy of the
// curve's pow and bisection, and no register of the row live across it, where
// it would spill), and loads them back; a lane touches only its own samples.
template <typename T, int NTH, int PER>
__device__ __forceinline__ void gain_step(const DevGain &G, T (&v)[PER], int n, double start,
                                          char *lds) {
#pragma clang fp contract(off)
  if (isnan(start)) return;
  const int tid_ = opaque_tid();
  const double dt = uni(G.dt_ms);
  const double u = (start - uni(G.t_min_ms)) / dt;  // the table position of sample 0
  int64_t s = 0;
  int j0 = n, j1 = n;
  bool on_grid = false;
  double f = 0;
  if (isfinite(u) && fabs(u) <= 1e15) {
    const double r = round(u);
    on_grid = fabs(u - r) <= 1e-9 * fmax(1.0, fabs(u));
    s = (int64_t)(on_grid ? r : floor(u));
    f = u - (double)s;
    const int64_t last = (int64_t)uni((int)G.len) - (on_grid ? 1 : 2);  // the last entry s + j may take
    const int64_t a = min(max(-s, (int64_t)0), (int64_t)n);
    const int64_t b = min(max(last - s + 1, a), (int64_t)n);
    j0 = (int)a;
    j1 = (int)b;
  }
  gptr<T> tab = as_global<T>(uni_ptr(G.table));
  const T ft = (T)f;
  T *x = (T *)lds;
  __syncthreads();  // the previous step's LDS reads are done
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int j = tid_ + NTH * k;
    if (j < n) x[j] = v[k];
  }
#pragma unroll 1
  for (int j = tid_; j < n; j += NTH) {
    const T xv = x[j];
    if (j >= j0 && j < j1) {
      const T a = tab[s + j];
      x[j] = xv * (on_grid ? a : a + ft * (tab[s + j + 1] - a));
    } else if (xv != (T)0) {
      x[j] = xv * (T)gain_factor_at(G, start + (double)j * dt);
    }
  }
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int j = tid_ + NTH * k;
    v[k] = j < n ? x[j] : (T)0;  // slots past the live end hold +0
  }
  __syncthreads();  // before the next step writes LDS
}

// One pass of one biquad section (sosKernel, filter/iir.nim:319-351) over
// u[0 ..< m] in place, forward or backward, as a parallel scan: thread t
// takes positions [t·C, t·C + C) of the pass (C odd, so that the lanes'
// float64 accesses spread over the LDS banks).  The transposed direct form
// II state s = (s1, s2) advances as s' = A·s + B·x with A = [[-a1, 1],
// [-a2, 0]], so a chunk is the affine map s -> M·s + e, M = A^len and e the
// chunk's end state from rest.  A steady-state start on the first input u0
// is the same as a start from rest on x − u0 plus the constant output g·u0,
// so the scan runs on x − u0 from rest (a constant pad then scans as exact
// zeros, and the rounding scales with the signal, not with the level it
// sits on).  (1) each thread runs its chunk from rest for e, and the two unit
// states with no input for M; (2) an inclusive scan of the maps over the
// threads: Hillis-Steele by shuffles inside each wave, then the wave totals
// through LDS (`sc`: six float64 per wave); (3) each thread runs its chunk
// again from its start state (the maps of the chunks before it applied to
// rest), writing the outputs.  The outputs equal the
// serial recursion's up to float64 rounding.
template <int NTH>
__device__ __forceinline__ void sos_scan_pass(double *u, int m, const double *c6, bool steady,
                                              bool backward, double *sc) {
  const int tid_ = opaque_tid();
  const double b0 = uni(c6[0]), b1 = uni(c6[1]), b2 = uni(c6[2]);
  const double a1 = uni(c6[3]), a2 = uni(c6[4]), g = uni(c6[5]);
  int C = (m + NTH - 1) / NTH;
  C |= 1;
  const int p0 = min(tid_ * C, m), p1 = min(p0 + C, m);
  // the level the pass starts from (steady state) or 0 (rest)
  const double u0 = steady ? u[backward ? m - 1 : 0] : 0.0;
  // (1) the chunk from rest, and A^len on the unit states: three independent
  // chains in one loop, so that each hides the others' latency
  double e1 = 0, e2 = 0;
  double m11 = 1, m21 = 0, m12 = 0, m22 = 1;  // columns: images of (1,0) and (0,1)
#pragma unroll 4
  for (int p = p0; p < p1; p++) {
    const int j = backward ? m - 1 - p : p;
    const double x = u[j] - u0;
    const double y = b0 * x + e1;
    const double t1 = b1 * x - a1 * y + e2;
    e2 = b2 * x - a2 * y;
    e1 = t1;
    const double n11 = -a1 * m11 + m21, n21 = -a2 * m11;
    const double n12 = -a1 * m12 + m22, n22 = -a2 * m12;
    m11 = n11; m21 = n21; m12 = n12; m22 = n22;
  }
  // (2) the scan: thread t composes the maps of chunks 0 .. t — inside each
  // wave by shuffles (no barrier), then across waves through the wave totals
  // in LDS (one barrier)
  const int lane = tid_ & 63, wid = tid_ >> 6;
  for (int d = 1; d < 64; d <<= 1) {
    const double q11 = __shfl_up(m11, d), q21 = __shfl_up(m21, d);
    const double q12 = __shfl_up(m12, d), q22 = __shfl_up(m22, d);
    const double f1 = __shfl_up(e1, d), f2 = __shfl_up(e2, d);
    if (lane >= d) {  // the chunks before apply first
      const double r11 = m11 * q11 + m12 * q21, r21 = m21 * q11 + m22 * q21;
      const double r12 = m11 * q12 + m12 * q22, r22 = m21 * q12 + m22 * q22;
      const double g1 = m11 * f1 + m12 * f2 + e1, g2 = m21 * f1 + m22 * f2 + e2;
      m11 = r11; m21 = r21; m12 = r12; m22 = r22; e1 = g1; e2 = g2;
    }
  }
  // the end state of the waves before this one, from rest
  double w1 = 0, w2 = 0;
  if constexpr (NTH > 64) {
    if (lane == 63) {
      double *S = sc + 6 * wid;
      S[0] = m11; S[1] = m21; S[2] = m12; S[3] = m22; S[4] = e1; S[5] = e2;
    }
    __syncthreads();
    for (int w = 0; w < wid; w++) {  // whole waves, in order
      const double *Q = sc + 6 * w;
      const double t1 = Q[0] * w1 + Q[2] * w2 + Q[4], t2 = Q[1] * w1 + Q[3] * w2 + Q[5];
      w1 = t1;
      w2 = t2;
    }
  }
  // (3) the chunk again from its start state: the end state of chunks
  // 0 .. t-1 from rest — the lane below's (its composition applied to the
  // waves before), or the waves before alone for a wave's first lane
  const double x1 = __shfl_up(m11 * w1 + m12 * w2 + e1, 1);
  const double x2 = __shfl_up(m21 * w1 + m22 * w2 + e2, 1);
  double s1 = lane > 0 ? x1 : w1, s2 = lane > 0 ? x2 : w2;
  const double y0 = g * u0;  // the steady output of the level
#pragma unroll 4
  for (int p = p0; p < p1; p++) {
    const int j = backward ? m - 1 - p : p;
    const double x = u[j] - u0;
    const double y = b0 * x + s1;
    s1 = b1 * x - a1 * y + s2;
    s2 = b2 * x - a2 * y;
    u[j] = y + y0;
  }
  __syncthreads();  // the pass is done (and `sc` free) before the next reads
}

// An IIR step (applyFilter, filter/iir.nim:363-433) on the row in registers:
// the trend removed (float64 block sums), the row with its pads (odd
// reflections before, and after for zero phase, or zeros after) as float64
// in LDS, each section's pass as a parallel scan (forward, then backward for
// zero phase), the trend's steady-state response G(1)·(a + b·i) − b·G'(1)
// added back on the way to the registers.
template <typename T, int NTH, int PER>
__device__ __forceinline__ void iir_step(const DevIir &I, T (&v)[PER], int n, char *lds,
                                         RedSlots &red) {
  const int tid_ = opaque_tid();
  double a, b;
  trend_of<T, NTH, PER>(v, n, uni(I.trend), red, a, b);
  a = uni(a);  // the same on every lane: SGPRs across the passes
  b = uni(b);
  const int lead = uni(I.lead), tail = uni(I.tail);
  const bool reflect = uni(I.reflect) != 0;
  double *u = (double *)lds;
  double *sc = u + uni(I.span);
  __syncthreads();  // the previous step's LDS reads are done
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int j = tid_ + NTH * k;
    if (j < n) u[lead + j] = (double)v[k] - (a + b * (double)j);
  }
  __syncthreads();
  // the pads read only the data
  if (reflect) {
    const double d0 = u[lead], dn = u[lead + n - 1];
    for (int k = tid_ + 1; k <= lead; k += NTH) u[lead - k] = 2 * d0 - u[lead + min(k, n - 1)];
    for (int k = tid_ + 1; k <= tail; k += NTH)
      u[lead + n - 1 + k] = 2 * dn - u[lead + n - 1 - min(k, n - 1)];
  } else {
    for (int k = tid_ + 1; k <= tail; k += NTH) u[lead + n - 1 + k] = 0;
  }
  __syncthreads();
  const int m = lead + n + tail;
  const int ns = uni(I.n_sections);
  const double *sec = (const double *)uni_ptr(I.sec);
  // the forward passes, then (zero phase) the backward ones: one loop, so
  // that the pass is inlined once
  const int passes = uni(I.zero_phase) ? 2 * ns : ns;
  for (int i = 0; i < passes; i++)
    sos_scan_pass<NTH>(u, m, sec + 6 * (i < ns ? i : i - ns), reflect, i >= ns, sc);
  const double ga = a * uni(I.dc_gain) - b * uni(I.dc_slope), gb = b * uni(I.dc_gain);
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int j = tid_ + NTH * k;
    v[k] = j < n ? (T)(u[lead + j] + (ga + gb * (double)j)) : (T)0;
  }
  __syncthreads();  // before the next step writes LDS
}

// ------------------------------------------ SEG-Y and SEG-D sample words

// IBM System/360 single precision: sign, 7-bit base-16 exponent in excess
// 64, 24-bit fraction below the radix point: (-1)^s · f/2^24 · 16^(e-64).
// The value is exact in float64 (24 significant bits, binary exponents
// -280 .. 228), so float32 output is one correctly rounded conversion.
__device__ __forceinline__ double ibm_to_double(uint32_t w) {
  const double mag = ldexp((double)(w & 0xFFFFFFu), 4 * (int)((w >> 24) & 0x7Fu) - 280);
  return (w >> 31) ? -mag : mag;
}

// Bytes per sample of a format: SEG-Y codes 1 (IBM), 2 (int32), 3 (int16),
// 5 (IEEE float32), 8 (int8); SEG-D demultiplexed codes 8036 (24-bit
// integer), 8038 (int32), 8058 (IEEE float32), 8080 (IEEE float64).
constexpr int sample_bytes(int fmt) {
  return fmt == 3 ? 2 : fmt == 8 ? 1 : fmt == 8036 ? 3 : fmt == 8080 ? 8 : 4;
}

// The raw word of sample j (big-endian, as a file holds it), widened; 8-byte
// words from two aligned 4-byte loads.
template <int FMT>
__device__ __forceinline__ uint64_t load_word(const unsigned char *p) {
  if constexpr (FMT == 3) return (uint64_t)((p[0] << 8) | p[1]);
  else if constexpr (FMT == 8) return (uint64_t)p[0];
  else if constexpr (FMT == 8036) return (uint64_t)((p[0] << 16) | (p[1] << 8) | p[2]);
  else if constexpr (FMT == 8080)
    return ((uint64_t)__builtin_bswap32(((const uint32_t *)p)[0]) << 32) |
           __builtin_bswap32(((const uint32_t *)p)[1]);
  else return __builtin_bswap32(*(const uint32_t *)p);
}

// The word's value as T: IBM through its exact float64 value; integers
// straight to T (exact but for int32 to float32, one rounding to nearest);
// IEEE words as they are (a NaN keeps its payload in its own width).
template <int FMT, typename T>
__device__ __forceinline__ T word_value(uint64_t w) {
  if constexpr (FMT == 1) return (T)ibm_to_double((uint32_t)w);
  else if constexpr (FMT == 2 || FMT == 8038) return (T)(int32_t)(uint32_t)w;
  else if constexpr (FMT == 3) return (T)(int16_t)(uint16_t)w;
  else if constexpr (FMT == 5 || FMT == 8058) return (T)__uint_as_float((uint32_t)w);
  else if constexpr (FMT == 8036) return (T)(((int32_t)((uint32_t)w << 8)) >> 8);
  else if constexpr (FMT == 8080) return (T)__longlong_as_double((long long)w);
  else return (T)(int8_t)(uint8_t)w;
}

// A raw word of format `fmt` at p as T, the format chosen at run time (one
// copy of each conversion, for loops that are not unrolled).
template <typename T>
__device__ __forceinline__ T raw_sample(int fmt, const unsigned char *p) {
  switch (fmt) {
    case 1: return word_value<1, T>(load_word<1>(p));
    case 2: return word_value<2, T>(load_word<2>(p));
    case 3: return word_value<3, T>(load_word<3>(p));
    case 5: return word_value<5, T>(load_word<5>(p));
    case 8036: return word_value<8036, T>(load_word<8036>(p));
    case 8038: return word_value<8038, T>(load_word<8038>(p));
    case 8058: return word_value<8058, T>(load_word<8058>(p));
    case 8080: return word_value<8080, T>(load_word<8080>(p));
    default: return word_value<8, T>(load_word<8>(p));
  }
}

__device__ __forceinline__ int raw_sample_bytes(int fmt) { return sample_bytes(fmt); }

// Per-trace normalization (gain/gain.nim:198-219).
template <typename T, int NTH, int PER>
__device__ __forceinline__ void norm_trace_step(const DevStep &st, T (&v)[PER], int n, char *lds,
                                RedSlots &red) {
  const int tid_ = opaque_tid();
  using S = typename Bits<T>::S;
  __shared__ S smax[(NTH / 64)];
  __shared__ int64_t sh_i[2];
  // max |x| as the integer maximum of the sign-cleared bits (a NaN sorts
  // above Inf: gain/stats.nim:97-117)
  S mb = 0;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    int j = tid_ + NTH * k;
    if (j < n) {
      S b = Bits<T>::abs_bits(v[k]);
      mb = b > mb ? b : mb;
    }
  }
  mb = wave_max_u<S>(mb);
  __syncthreads();
  if ((tid_ & 63) == 0) smax[tid_ >> 6] = mb;
  __syncthreads();
  mb = smax[0];
  for (int i = 1; i < (NTH / 64); i++) mb = smax[i] > mb ? smax[i] : mb;
  const double mx = Bits<T>::value(mb);
  double level;
  const int stat = uni(st.stat);
  const double starget = uni(st.target);
  if (stat == 0) {
    level = mx;
  } else if (stat == 1) {
    double sq = 0;
#pragma unroll
    for (int k = 0; k < PER; k++) {
      int j = tid_ + NTH * k;
      if (j < n) sq += (double)v[k] * (double)v[k];
    }
    level = sqrt(block_sum<NTH>(sq, red.next()) / (double)n);
  } else {
    // percentileAbs (gain/stats.nim:159-182): linear between order
    // statistics k and k + 1 of |x|, h = (n - 1) p / 100.
    if (isnan(mx)) return;
    double hh = (double)(n - 1) * uni(st.percentile) / 100;
    int64_t k = (int64_t)floor(hh);
    if (k > n - 1) k = n - 1;
    uint32_t *hist = (uint32_t *)lds;
    S lo = select_bits<T, NTH, PER>(v, n, k, hist, sh_i);
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
        int j = tid_ + NTH * q;
        if (j < n) {
          S b = Bits<T>::abs_bits(v[q]);
          if (b <= lo) cnt += 1;
          else up = b < up ? b : up;
        }
      }
      cnt = block_sum<NTH>(cnt, red.next());
      // block minimum of up
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        S w2 = __shfl_xor(up, o, 64);
        up = w2 < up ? w2 : up;
      }
      __syncthreads();
      if ((tid_ & 63) == 0) smax[tid_ >> 6] = up;
      __syncthreads();
      up = smax[0];
      for (int i = 1; i < (NTH / 64); i++) up = smax[i] < up ? smax[i] : up;
      double upper = cnt >= (double)(k + 2) ? lower : Bits<T>::value(up);
      double f = hh - (double)k;
      level = f == 0 ? lower : lower + f * (upper - lower);
    }
  }
  // scaleToTarget (gain/gain.nim:208-210)
  if (isfinite(level) && level > 0 && fits_in<T>(starget / level)) {
    const T f = (T)(starget / level);
#pragma unroll
    for (int k = 0; k < PER; k++) {
      int j = tid_ + NTH * k;
      if (j < n) v[k] = v[k] * f;
    }
  }
  __syncthreads();
}

// Clears n uint32 (a kernel, not hipMemsetAsync, so that a captured graph
// holds no memset node).
__global__ void zero_u32_kernel(uint32_t *__restrict__ p, int32_t n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

}  // namespace

}  // namespace pseg
