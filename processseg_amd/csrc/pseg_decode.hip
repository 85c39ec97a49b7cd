// SEG-Y and SEG-D sample decode on the device: rows of big-endian samples as
// a file holds them (each trace a header then its samples, `pitch` bytes
// apart) to a [rows, cols] float32 or float64 panel, one pass over HBM.
// Formats: SEG-Y rev 2 data sample format codes 1 IBM float32, 2 int32,
// 3 int16, 5 IEEE float32, 8 int8; SEG-D demultiplexed codes 8036 24-bit
// integer, 8038 int32, 8058 IEEE float32, 8080 IEEE float64 — what
// parseSEG's decode produces before the prefetcher's transform runs the plan
// (batch.nim:267-309).
#include "pseg_device.hpp"

namespace pseg {

namespace {

// One workgroup of 256 threads per row; each thread loads K words a round
// before it converts and stores them (K loads in flight).  Measured on the
// 2000 x 3001 IBM panel: K = 4 12.2 us per launch, K = 12 14.4 us.
constexpr int K = 4;
template <int FMT, typename T>
__global__ __launch_bounds__(256) void decode_kernel(const unsigned char *__restrict__ raw,
                                                     int64_t pitch, int64_t offset, int32_t cols,
                                                     T *__restrict__ dst) {
  constexpr int B = sample_bytes(FMT);
  const int64_t r = blockIdx.x;
  const unsigned char *row = raw + r * pitch + offset;
  T *out = dst + r * (int64_t)cols;
  for (int j0 = (int)threadIdx.x; j0 < cols; j0 += K * 256) {
    uint64_t w[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int j = j0 + 256 * k;
      w[k] = j < cols ? load_word<FMT>(row + (int64_t)j * B) : 0u;
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
      const int j = j0 + 256 * k;
      if (j < cols) out[j] = word_value<FMT, T>(w[k]);
    }
  }
}

template <typename T>
hipError_t launch_fmt(int format, const void *raw, int64_t pitch, int64_t offset, int32_t rows,
                      int32_t cols, void *dst, hipStream_t s) {
  const unsigned char *p = (const unsigned char *)raw;
  T *d = (T *)dst;
  switch (format) {
    case 1: hipLaunchKernelGGL((decode_kernel<1, T>), dim3(rows), dim3(256), 0, s, p, pitch, offset, cols, d); break;
    case 2: hipLaunchKernelGGL((decode_kernel<2, T>), dim3(rows), dim3(256), 0, s, p, pitch, offset, cols, d); break;
    case 3: hipLaunchKernelGGL((decode_kernel<3, T>), dim3(rows), dim3(256), 0, s, p, pitch, offset, cols, d); break;
    case 5: hipLaunchKernelGGL((decode_kernel<5, T>), dim3(rows), dim3(256), 0, s, p, pitch, offset, cols, d); break;
    case 8: hipLaunchKernelGGL((decode_kernel<8, T>), dim3(rows), dim3(256), 0, s, p, pitch, offset, cols, d); break;
    case 8036: hipLaunchKernelGGL((decode_kernel<8036, T>), dim3(rows), dim3(256), 0, s, p, pitch, offset, cols, d); break;
    case 8038: hipLaunchKernelGGL((decode_kernel<8038, T>), dim3(rows), dim3(256), 0, s, p, pitch, offset, cols, d); break;
    case 8058: hipLaunchKernelGGL((decode_kernel<8058, T>), dim3(rows), dim3(256), 0, s, p, pitch, offset, cols, d); break;
    case 8080: hipLaunchKernelGGL((decode_kernel<8080, T>), dim3(rows), dim3(256), 0, s, p, pitch, offset, cols, d); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace

hipError_t launch_decode(int dtype, int format, const void *raw, int64_t pitch, int64_t offset,
                         int32_t rows, int32_t cols, void *dst, hipStream_t s) {
  if (rows == 0 || cols == 0) return hipSuccess;
  return dtype == 0 ? launch_fmt<float>(format, raw, pitch, offset, rows, cols, dst, s)
                           : launch_fmt<double>(format, raw, pitch, offset, rows, cols, dst, s);
}

}  // namespace pseg
