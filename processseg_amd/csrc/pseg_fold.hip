// Fold counts (qc/fold.nim addFold) on the GPU: one thread per trace finds
// the trace's bin (stack/binning.nim) and counts it.  Every count is an
// integer atomic and the offset range a compare-and-swap minimum and
// maximum, so the map is the same whatever order the threads run in.  The
// bin arithmetic is the serial reference's, operation for operation, with
// no multiply-add contracted.
#pragma clang fp contract(off)

#include "pseg_internal.hpp"

#include <math.h>

namespace pseg {

namespace {

// nearest (stack/binning.nim:197-206): bin i holds [i - 0.5, i + 0.5); -1
// outside [-0.5, n - 0.5) or NaN.
__device__ __forceinline__ int64_t nearest(double u, int64_t n) {
  if (!(u >= -0.5 && u < (double)n - 0.5)) return -1;
  const double whole = floor(u);
  const int64_t k = (int64_t)whole + (u - whole >= 0.5 ? 1 : 0);
  return k < n - 1 ? k : n - 1;
}

__device__ __forceinline__ int64_t grid_bin(const pseg_binning &b, double u, double v) {
  const int64_t i = nearest(v, b.inlines), j = nearest(u, b.crosslines);
  return i < 0 || j < 0 ? -1 : i * b.crosslines + j;
}

// binAt (stack/binning.nim:213-226).
__device__ int64_t bin_at(const pseg_binning &b, double x, double y) {
  const double dx = x - b.x0, dy = y - b.y0;
  if (b.kind == PSEG_BIN_LINE) {
    if (!(fabs(dx * b.uy - dy * b.ux) <= b.max_off)) return -1;
    return nearest((dx * b.ux + dy * b.uy) / b.size0, b.crosslines);
  }
  if (b.kind == PSEG_BIN_GRID)
    return grid_bin(b, (dx * b.ux + dy * b.uy) / b.size0, (dx * b.vx + dy * b.vy) / b.size1);
  return -1;
}

// Atomic minimum and maximum of a float that starts NaN (empty): the
// result is the extreme of the values added, in any order.
__device__ void atomic_min_f32(float *p, float f) {
  unsigned int *q = (unsigned int *)p;
  unsigned int old = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (!(__uint_as_float(old) <= f)) {
    const unsigned int prev = atomicCAS(q, old, __float_as_uint(f));
    if (prev == old) break;
    old = prev;
  }
}

__device__ void atomic_max_f32(float *p, float f) {
  unsigned int *q = (unsigned int *)p;
  unsigned int old = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (!(__uint_as_float(old) >= f)) {
    const unsigned int prev = atomicCAS(q, old, __float_as_uint(f));
    if (prev == old) break;
    old = prev;
  }
}

// floorMod on floats (Nim's std/math): fmod, moved to the sign of y.
__device__ __forceinline__ double floor_mod(double x, double y) {
  double r = fmod(x, y);
  if ((r > 0 && y < 0) || (r < 0 && y > 0)) r += y;
  return r;
}

constexpr double kPi = 3.141592653589793;

// addFold (qc/fold.nim:583-588) with addTrace (:561-581) and addToBin
// (:429-451): one thread per row; the block's counters are summed in LDS
// and added once.
__global__ __launch_bounds__(256) void fold_add_kernel(
    pseg_binning b, const double *__restrict__ bounds, int32_t nb, int32_t sectors,
    double start_deg, int32_t offset_range, const double *__restrict__ keys,
    const int32_t *__restrict__ row_len, int32_t rows, int32_t *fold, int32_t *ofold,
    int32_t *afold, float *omin, float *omax, unsigned long long *counters) {
  __shared__ unsigned int cnt[4];
  if (threadIdx.x < 4) cnt[threadIdx.x] = 0;
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t size = b.inlines * b.crosslines;
  if (r < rows) {
    int what = -1;  // 0 binned, 1 outside, 2 no coordinates, 3 not live
    if (row_len[r] < 0) {
      what = 3;
    } else {
      const double *k = keys + r * PSEG_FOLD_KEYS;
      const double sx = k[PSEG_KEY_SRC_X], sy = k[PSEG_KEY_SRC_Y];
      const double rx = k[PSEG_KEY_RCV_X], ry = k[PSEG_KEY_RCV_Y];
      const bool coords = isfinite(sx) && isfinite(sy) && isfinite(rx) && isfinite(ry);
      int64_t bin = -1;
      if (b.kind == PSEG_BIN_LINE || b.kind == PSEG_BIN_GRID) {
        double x = k[PSEG_KEY_CMP_X], y = k[PSEG_KEY_CMP_Y];
        if (!(isfinite(x) && isfinite(y))) {
          if (!coords) {
            what = 2;
          } else {
            x = (sx + rx) * 0.5;
            y = (sy + ry) * 0.5;
          }
        }
        if (what < 0) bin = bin_at(b, x, y);
      } else if (b.kind == PSEG_BIN_CDP) {
        bin = nearest((k[PSEG_KEY_ENSEMBLE] - b.first0) / b.step0, b.crosslines);
      } else {
        bin = grid_bin(b, (k[PSEG_KEY_CROSSLINE] - b.first1) / b.step1,
                       (k[PSEG_KEY_INLINE] - b.first0) / b.step0);
      }
      if (what < 0 && bin < 0) what = 1;
      if (what < 0) {
        what = 0;
        const double offset = coords ? hypot(rx - sx, ry - sy) : fabs(k[PSEG_KEY_OFFSET]);
        atomicAdd(fold + bin, 1);
        if (nb > 1 && offset >= bounds[0] && offset < bounds[nb - 1]) {
          int32_t c = 0;  // bounds[c] <= offset < bounds[c + 1]
          while (c + 1 < nb && bounds[c + 1] <= offset) c++;
          atomicAdd(ofold + (int64_t)c * size + bin, 1);
        }
        if (offset_range && !isnan(offset)) {
          const float f = (float)offset;
          atomic_min_f32(omin + bin, f);
          atomic_max_f32(omax + bin, f);
        }
        if (sectors > 0 && coords) {
          // bearingOf (qc/fold.nim:453-455): radToDeg(arctan2(...)), the
          // degrees as d / (pi / 180)
          const double bearing = atan2(rx - sx, ry - sy) / (kPi / 180.0);
          if (!isnan(bearing)) {
            int64_t s = (int64_t)(floor_mod(bearing - start_deg, 360.0) * (double)sectors / 360.0);
            s = s < sectors - 1 ? s : sectors - 1;
            atomicAdd(afold + s * size + bin, 1);
          }
        }
      }
    }
    atomicAdd(&cnt[what], 1u);
  }
  __syncthreads();
  if (threadIdx.x < 4 && cnt[threadIdx.x])
    atomicAdd(counters + threadIdx.x, (unsigned long long)cnt[threadIdx.x]);
}

// clear (qc/fold.nim:416-427): counts to 0, the offset range to NaN.
__global__ void fold_clear_kernel(int64_t bins, int32_t classes, int32_t sectors,
                                  int32_t offset_range, int32_t *fold, int32_t *ofold,
                                  int32_t *afold, float *omin, float *omax,
                                  unsigned long long *counters) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int64_t i = t; i < bins; i += stride) {
    fold[i] = 0;
    if (offset_range) {
      omin[i] = __uint_as_float(0x7fc00000u);
      omax[i] = __uint_as_float(0x7fc00000u);
    }
  }
  for (int64_t i = t; i < bins * classes; i += stride) ofold[i] = 0;
  for (int64_t i = t; i < bins * sectors; i += stride) afold[i] = 0;
  if (t < 4) counters[t] = 0;
}

// foldDifference (qc/fold.nim:612-622): pre-plot minus actual, wrapping.
__global__ void fold_diff_kernel(const int32_t *__restrict__ pre, const int32_t *__restrict__ act,
                                 int32_t *__restrict__ out, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = (int32_t)((uint32_t)pre[i] - (uint32_t)act[i]);
}

int grid_for(int64_t n) {
  const int64_t g = (n + 255) / 256;
  return (int)(g < 1 ? 1 : g > 4096 ? 4096 : g);
}

}  // namespace

hipError_t launch_fold_add(const pseg_binning &b, const double *bounds, int32_t nb,
                           int32_t sectors, double start_deg, int32_t offset_range,
                           const double *keys, const int32_t *row_len, int32_t rows, int32_t *fold,
                           int32_t *ofold, int32_t *afold, float *omin, float *omax,
                           unsigned long long *counters, hipStream_t stream) {
  if (rows == 0) return hipSuccess;
  hipLaunchKernelGGL(fold_add_kernel, dim3((rows + 255) / 256), dim3(256), 0, stream, b, bounds,
                     nb, sectors, start_deg, offset_range, keys, row_len, rows, fold, ofold, afold,
                     omin, omax, counters);
  return hipGetLastError();
}

hipError_t launch_fold_clear(int64_t bins, int32_t classes, int32_t sectors, int32_t offset_range,
                             int32_t *fold, int32_t *ofold, int32_t *afold, float *omin,
                             float *omax, unsigned long long *counters, hipStream_t stream) {
  const int64_t most = bins * (int64_t)(classes > sectors ? (classes > 1 ? classes : 1)
                                                          : (sectors > 1 ? sectors : 1));
  hipLaunchKernelGGL(fold_clear_kernel, dim3(grid_for(most)), dim3(256), 0, stream, bins, classes,
                     sectors, offset_range, fold, ofold, afold, omin, omax, counters);
  return hipGetLastError();
}

hipError_t launch_fold_diff(const int32_t *pre, const int32_t *act, int32_t *out, int64_t n,
                            hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(fold_diff_kernel, dim3(grid_for(n)), dim3(256), 0, stream, pre, act, out, n);
  return hipGetLastError();
}

}  // namespace pseg
