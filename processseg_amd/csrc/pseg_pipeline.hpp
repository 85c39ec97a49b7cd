// The fused row kernel of the panel pipeline, templated on the row width:
// NTH threads (NTH / 64 wave64) per row, PER samples per thread; FORM 0:
// demean / filter / AGC only (less code for the instruction cache, the form
// most flows take), 1: every step kind but the IIR cascades, 2: every step
// kind (the IIR pass's float64 state costs the other steps a spill, so only
// flows with an IIR step take it).
//
// One workgroup per row.  The row crosses HBM once: it is loaded into
// registers (sample j on thread j mod NTH, slot j / NTH, so every load and
// store is coalesced), every step of the plan runs on it there and in LDS,
// and it is stored once.  Steps:
//   demean / detrend / scale / clip / polarity   core/basic.nim:52-110
//   FFT filter (trend, pads, rFFT, gain, irFFT)  filter/apply.nim:194-363
//   AGC (block prefix/suffix sums, gain step)     gain/gain.nim:351-461
//   per-trace normalization (max, rms, exact percentile by radix select)
//                                                 gain/gain.nim:198-219, gain/stats.nim:159-182
// Each pipeline_nt*.hip instantiates one NTH with its PER widths.
#pragma once
#include "pseg_device.hpp"

#ifndef PSEG_F32_WAVES
#define PSEG_F32_WAVES 4
#endif

#include <algorithm>

namespace pseg {

namespace {

// ---------------------------------------------------------------- kernels

// One row of one job: load the row's window into registers, run the job's
// steps on its live part, store the job's crop.  Every thread of the block
// runs it with the same (job, row).
template <typename T, int NTH, int PER, int FORM>
__device__ __forceinline__ void run_row(const Job &J, int64_t r, char *lds, RedSlots &red) {
  const int tid_ = opaque_tid();
  const DevPlan *plan = J.plan;
  const int cols = J.cols;
  // the row's samples through global (not flat) pointers: the loads and
  // stores take the row base in SGPRs and a 32-bit lane offset, and do not
  // count against the LDS wait counter
  gptr<T> s = as_global<T>((const T *)J.src + r * J.src_pitch);
  gmut<T> d = as_global_mut<T>((T *)J.dst + r * J.dst_pitch);
  const int num = as_global<int32_t>(J.num)[r];
  const int n_row = num < 0 ? -1 : min(max(num - J.first, 0), cols);
  // A calibration step scales every row's live part, dead and auxiliary
  // rows included (pipeline.nim:44-45): n_all, the samples any step may
  // change, is that length on a plan with one, else the live length.
  int n_all = n_row;
  if constexpr (FORM >= 1) {
    if (J.cal_len != nullptr) {
      const int cl = as_global<int32_t>(J.cal_len)[r];
      n_all = max(n_row, min(max(cl - J.first, 0), cols));
    }
  }
  const int n = n_all;
  T v[PER];
  bool raw = false;
  if constexpr (FORM >= 1) raw = uni(J.src_fmt) != 0;
  if (raw) {
    // raw SEG-Y traces: each lane decodes its own samples into LDS (one copy
    // of the conversions, in a loop that is not unrolled), then loads them
    const int fmt = uni(J.src_fmt);
    const int bytes = raw_sample_bytes(fmt);
    const unsigned char *rb = (const unsigned char *)J.src + r * J.src_pitch + J.src_off;
    T *st = (T *)lds;
#pragma unroll 1
    for (int j = tid_; j < cols; j += NTH) st[j] = raw_sample<T>(fmt, rb + (int64_t)j * bytes);
#pragma unroll
    for (int k = 0; k < PER; k++) {
      int j = tid_ + NTH * k;
      v[k] = j < cols ? st[j] : (T)0;
    }
    __syncthreads();  // before the steps write LDS
  } else {
#pragma unroll
    for (int k = 0; k < PER; k++) {
      int j = tid_ + NTH * k;
      v[k] = j < cols ? s[j] : (T)0;
    }
  }
  // Samples past the live end leave now, unchanged: every step changes live
  // samples only.  Their slots then hold +0 for the steps (sums over every
  // slot need no test of the live length; steps keep them +0).
  const int c0 = J.crop0, c1 = J.crop1;
  if (J.src != J.dst || c0 != 0) {
    const int live = max(n, 0);
    gmut<T> dd = d - c0;
#pragma unroll
    for (int k = 0; k < PER; k++) {
      int j = tid_ + NTH * k;
      if (j >= live && j >= c0 && j < c1) dd[j] = v[k];
    }
  }
#pragma unroll
  for (int k = 0; k < PER; k++) {
    int j = tid_ + NTH * k;
    v[k] = j < n ? v[k] : (T)0;
  }
  if (n > 0) {
    if (J.has_factor && n_row > 0) {
      // a per-panel factor at the start of the part (pipeline.nim:717-718)
      const double f = *J.factor;
      if (!isnan(f)) {
        const T kf = (T)f;
#pragma unroll
        for (int k = 0; k < PER; k++) {
          int j = tid_ + NTH * k;
          if (j < n) v[k] = v[k] * kf;
        }
      }
    }
    for (int i = J.step_first; i < J.step_last; i++) {
      const DevStep &st = plan->steps[i];
      if constexpr (FORM >= 1) {
        if (uni(st.kind) == kCalibrate) {
          // calibrate / uncalibrate (core/calibrate.nim:582-593, 615-624):
          // the row's live part times its multiplier in T; 1 keeps the row
          const double kd = uni(as_global<double>(J.cal_k)[(int64_t)uni(st.stat) * J.rows + r]);
          if (kd != 1.0) {
            const int tid_ = opaque_tid();
            const int nc = opaque_s(uni(n_all));
            const T kk = (T)kd;
#pragma unroll
            for (int k = 0; k < PER; k++) {
              int j = tid_ + NTH * k;
              if (j < nc) v[k] = v[k] * kk;
            }
          }
          continue;
        }
        if (n_row <= 0) continue;  // every other step: live rows only
      }
      const int n = opaque_s(uni(n_row));
      switch (uni(st.kind)) {
        case kDemean: {
          const int tid_ = opaque_tid();
          // demean (core/basic.nim:52-59)
          const T m = (T)(block_sum<NTH>(row_sum<T, PER>(v), red.next()) / (double)n);
#pragma unroll
          for (int k = 0; k < PER; k++) {
            int j = tid_ + NTH * k;
            if (j < n) v[k] = v[k] - m;
          }
          break;
        }
        case kFilter: filter_step<T, NTH, PER>(st.filter, v, n, lds, red); break;
        case kAgc: agc_step<T, NTH, PER>(st.agc, v, n, lds, red); break;
        case kDetrend:
          if constexpr (FORM >= 1) {
            const int tid_ = opaque_tid();
            // detrend (core/basic.nim:61-81)
            double a, b;
            trend_of<T, NTH, PER>(v, n, 2, red, a, b);
            const T ta = (T)a, tb = (T)b;
#pragma unroll
            for (int k = 0; k < PER; k++) {
              int j = tid_ + NTH * k;
              if (j < n) v[k] = v[k] - (ta + tb * (T)j);
            }
          }
          break;
        case kScale:
          if constexpr (FORM >= 1) {
            const int tid_ = opaque_tid();
            const T kf = (T)uni(st.factor);
#pragma unroll
            for (int k = 0; k < PER; k++) {
              int j = tid_ + NTH * k;
              if (j < n) v[k] = v[k] * kf;
            }
          }
          break;
        case kClip:
          if constexpr (FORM >= 1) {
            const int tid_ = opaque_tid();
            const T lo = (T)uni(st.lo), hi = (T)uni(st.hi);
#pragma unroll
            for (int k = 0; k < PER; k++) {
              int j = tid_ + NTH * k;
              if (j < n) {
                T x = v[k];
                v[k] = x > hi ? hi : (x < lo ? lo : x);
              }
            }
          }
          break;
        case kPolarity:
          if constexpr (FORM >= 1) {
            const int tid_ = opaque_tid();
#pragma unroll
            for (int k = 0; k < PER; k++) {
              int j = tid_ + NTH * k;
              if (j < n) v[k] = -v[k];
            }
          }
          break;
        case kNormTrace:
          if constexpr (FORM >= 1) norm_trace_step<T, NTH, PER>(st, v, n, lds, red);
          break;
        case kMuteTop:
        case kMuteBottom:
          if constexpr (FORM >= 1) {
            // muteCol (core/mute.nim:189-194) from the row's keys (the
            // job's fields through readfirstlane: SGPRs, whatever the
            // compiler hoists)
            const double *keys = (const double *)uni_ptr(J.keys);
            const double col0 = uni((double)J.col0);
            const double t0 = keys ? uni(keys[2 * r]) : 0.0;
            const double off = keys ? uni(keys[2 * r + 1]) : (double)NAN;
            const double col = (mute_time_ms(st.mute, off) - t0) / uni(st.mute.dt_ms) - col0;
            mute_step<T, NTH, PER>(st.mute, uni(st.kind) == kMuteTop, v, n, col);
          }
          break;
        case kIir:
          if constexpr (FORM == 2) iir_step<T, NTH, PER>(st.iir, v, n, lds, red);
          break;
        case kGainCurve:
          if constexpr (FORM >= 1) {
            // sampleTimeMs(r, 0) = t0Ms + firstSample·dt (core/types.nim:237-240)
            const double *keys = (const double *)uni_ptr(J.keys);
            const double t0 = keys ? uni(keys[2 * r]) : 0.0;
            const double start = t0 + uni((double)J.col0) * uni(st.gain.dt_ms);
            gain_step<T, NTH, PER>(st.gain, v, n, start, lds);
          }
          break;
        default: break;
      }
    }
  }
  // the live samples
  {
    const int tid_ = opaque_tid();
    gmut<T> dd = d - c0;
#pragma unroll
    for (int k = 0; k < PER; k++) {
      int j = tid_ + NTH * k;
      if (j < n && j >= c0 && j < c1) dd[j] = v[k];
    }
  }
}

// The job holding work item `item`: the last with item0 <= item.
__device__ __forceinline__ int find_job(const Job *__restrict__ jobs, int n_jobs, int item) {
  int lo = 0, hi = n_jobs - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].item0 <= item) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

}  // namespace

// The pipeline over work items (one item = one row of one job).  Direct:
// item = blockIdx.x.  Persistent (counter != NULL): every block takes items
// from the counter until they run out, so rows of very different cost
// balance across the chip; each block exits once the counter passes `total`.
// n_jobs == 0: the single job passed by value.
template <typename T, int NTH, int PER, int FORM>
__global__ __launch_bounds__(NTH, sizeof(T) == 4 ? PSEG_F32_WAVES : 1) void pipeline_kernel(Job single, const Job *__restrict__ jobs,
                                                      int32_t n_jobs, int32_t total,
                                                      uint32_t *__restrict__ counter) {
  extern __shared__ __align__(16) char lds[];
  __shared__ double red_lds[4 * (NTH / 64)];
  __shared__ int s_item;
  int item = blockIdx.x;
  for (;;) {
    if (counter != nullptr) {
      if (threadIdx.x == 0) s_item = (int)atomicAdd(counter, 1u);
      __syncthreads();
      item = __builtin_amdgcn_readfirstlane(s_item);
      __syncthreads();
    }
    if (item >= total) break;
    int64_t row = item;
    const Job *jp = &single;
    Job J;
    if (n_jobs == 0) {
      J = single;
    } else {
      const int j = __builtin_amdgcn_readfirstlane(find_job(jobs, n_jobs, item));
      const Job &G = jobs[j];
      J.plan = uni_ptr(G.plan);
      J.src = uni_ptr(G.src);
      J.dst = uni_ptr(G.dst);
      J.num = uni_ptr(G.num);
      J.factor = uni_ptr(G.factor);
      J.keys = uni_ptr(G.keys);
      J.col0 = ((int64_t)uni((int)(G.col0 >> 32)) << 32) | (uint32_t)uni((int)G.col0);
      J.src_off = ((int64_t)uni((int)(G.src_off >> 32)) << 32) | (uint32_t)uni((int)G.src_off);
      J.src_fmt = uni(G.src_fmt);
      J.src_pitch = ((int64_t)uni((int)(G.src_pitch >> 32)) << 32) | (uint32_t)uni((int)G.src_pitch);
      J.dst_pitch = ((int64_t)uni((int)(G.dst_pitch >> 32)) << 32) | (uint32_t)uni((int)G.dst_pitch);
      J.rows = uni(G.rows);
      J.cols = uni(G.cols);
      J.first = uni(G.first);
      J.crop0 = uni(G.crop0);
      J.crop1 = uni(G.crop1);
      J.item0 = uni(G.item0);
      J.step_first = uni(G.step_first);
      J.step_last = uni(G.step_last);
      J.has_factor = uni(G.has_factor);
      J.cal_len = uni_ptr(G.cal_len);
      J.cal_k = uni_ptr(G.cal_k);
      row = item - J.item0;
    }
    (void)jp;
    RedSlots red{red_lds, 2 * (NTH / 64), 0};
    run_row<T, NTH, PER, FORM>(J, row, lds, red);
    if (counter == nullptr) break;
  }
}

namespace {

// Lets the kernel take up to kMaxRowLds bytes of dynamic LDS (once).
template <typename T, int NTH, int PER, int FORM>
hipError_t prepare_kernel() {
  static const hipError_t e = hipFuncSetAttribute(
      (const void *)pipeline_kernel<T, NTH, PER, FORM>, hipFuncAttributeMaxDynamicSharedMemorySize,
      (int)kMaxRowLds);
  return e;
}

// One launch of pipeline_kernel<T, NTH, PER, FORM>: clears the work counter of a
// persistent launch first (a kernel, not a memset, so the launch can be
// captured in a graph).
template <typename T, int NTH, int PER, int FORM>
hipError_t launch_one(const JobLaunch &a, hipStream_t stream) {
  hipError_t e = prepare_kernel<T, NTH, PER, FORM>();
  if (e != hipSuccess) return e;
  if (a.counter) {
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(64), 0, stream, a.counter, 1);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  const int grid = a.counter ? a.grid : a.total;
  hipLaunchKernelGGL((pipeline_kernel<T, NTH, PER, FORM>), dim3(grid), dim3(NTH), a.lds_bytes, stream,
                     a.single, a.jobs, a.n_jobs, a.total, a.counter);
  return hipGetLastError();
}

// Workgroups of pipeline_kernel<T, NTH, PER, FORM> resident on the whole device at
// once with `lds` bytes of dynamic LDS each (the grid of a persistent launch).
template <typename T, int NTH, int PER, int FORM>
hipError_t resident_one(size_t lds, int *blocks) {
  hipError_t e = prepare_kernel<T, NTH, PER, FORM>();
  if (e != hipSuccess) return e;
  int per_cu = 0, dev = 0, cus = 0;
  e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, (const void *)pipeline_kernel<T, NTH, PER, FORM>, NTH, lds);
  if (e != hipSuccess) return e;
  e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (e != hipSuccess) return e;
  *blocks = std::max(per_cu, 1) * cus;
  return hipSuccess;
}

}  // namespace

}  // namespace pseg
