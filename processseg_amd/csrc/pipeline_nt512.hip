// pipeline_kernel with 512 threads (8 wave64) per row: rows of up to
// 12288 samples, 8, 16, 24 samples per thread.
#include "pseg_pipeline.hpp"

namespace pseg {

namespace {

template <typename T>
hipError_t launch_widths(const JobLaunch &a, hipStream_t stream) {
  const int per = (a.max_cols + 511) / 512;
  if (per <= 8) return launch_one<T, 512, 8>(a, stream);
  if (per <= 16) return launch_one<T, 512, 16>(a, stream);
  if (per <= 24) return launch_one<T, 512, 24>(a, stream);
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t resident_widths(int max_cols, size_t lds, int *blocks) {
  const int per = (max_cols + 511) / 512;
  if (per <= 8) return resident_one<T, 512, 8>(lds, blocks);
  if (per <= 16) return resident_one<T, 512, 16>(lds, blocks);
  if (per <= 24) return resident_one<T, 512, 24>(lds, blocks);
  return hipErrorInvalidValue;
}

}  // namespace

hipError_t launch_pipeline_nt512(int dtype, const JobLaunch &a, hipStream_t stream) {
  return dtype == 0 ? launch_widths<float>(a, stream) : launch_widths<double>(a, stream);
}

hipError_t resident_pipeline_nt512(int dtype, int max_cols, size_t lds, int *blocks) {
  return dtype == 0 ? resident_widths<float>(max_cols, lds, blocks)
                    : resident_widths<double>(max_cols, lds, blocks);
}

}  // namespace pseg
