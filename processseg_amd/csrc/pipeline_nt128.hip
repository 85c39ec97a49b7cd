// pipeline_kernel with 128 threads (2 wave64) per row: rows of up to
// 3072 samples, 4, 8, 12, 16, 24 samples per thread.
#include "pseg_pipeline.hpp"

namespace pseg {

namespace {

template <typename T>
hipError_t launch_widths(const JobLaunch &a, hipStream_t stream) {
  const int per = (a.max_cols + 127) / 128;
  if (per <= 4) return launch_one<T, 128, 4>(a, stream);
  if (per <= 8) return launch_one<T, 128, 8>(a, stream);
  if (per <= 12) return launch_one<T, 128, 12>(a, stream);
  if (per <= 16) return launch_one<T, 128, 16>(a, stream);
  if (per <= 24) return launch_one<T, 128, 24>(a, stream);
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t resident_widths(int max_cols, size_t lds, int *blocks) {
  const int per = (max_cols + 127) / 128;
  if (per <= 4) return resident_one<T, 128, 4>(lds, blocks);
  if (per <= 8) return resident_one<T, 128, 8>(lds, blocks);
  if (per <= 12) return resident_one<T, 128, 12>(lds, blocks);
  if (per <= 16) return resident_one<T, 128, 16>(lds, blocks);
  if (per <= 24) return resident_one<T, 128, 24>(lds, blocks);
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_pipeline_nt128(int dtype, const JobLaunch &a, hipStream_t stream) {
  return dtype == 0 ? launch_widths<float>(a, stream) : launch_widths<double>(a, stream);
}

hipError_t resident_pipeline_nt128(int dtype, int max_cols, size_t lds, int *blocks) {
  return dtype == 0 ? resident_widths<float>(max_cols, lds, blocks)
                    : resident_widths<double>(max_cols, lds, blocks);
}

}  // namespace pseg
