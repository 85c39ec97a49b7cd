// pipeline_kernel with 128 threads (2 wave64) per row: rows of up to
// 3072 samples, 4, 8, 12, 16, 24 samples per thread; each width in
// the demean / filter / AGC form and the form with every step.
#include "pseg_pipeline.hpp"

namespace pseg {

namespace {

template <typename T, bool FULL>
hipError_t launch_widths(const JobLaunch &a, hipStream_t stream) {
  const int per = (a.max_cols + 127) / 128;
  if (per <= 4) return launch_one<T, 128, 4, FULL>(a, stream);
  if (per <= 8) return launch_one<T, 128, 8, FULL>(a, stream);
  if (per <= 12) return launch_one<T, 128, 12, FULL>(a, stream);
  if (per <= 16) return launch_one<T, 128, 16, FULL>(a, stream);
  if (per <= 24) return launch_one<T, 128, 24, FULL>(a, stream);
  return hipErrorInvalidValue;
}

template <typename T, bool FULL>
hipError_t resident_widths(int max_cols, size_t lds, int *blocks) {
  const int per = (max_cols + 127) / 128;
  if (per <= 4) return resident_one<T, 128, 4, FULL>(lds, blocks);
  if (per <= 8) return resident_one<T, 128, 8, FULL>(lds, blocks);
  if (per <= 12) return resident_one<T, 128, 12, FULL>(lds, blocks);
  if (per <= 16) return resident_one<T, 128, 16, FULL>(lds, blocks);
  if (per <= 24) return resident_one<T, 128, 24, FULL>(lds, blocks);
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_forms(const JobLaunch &a, hipStream_t stream) {
  return a.full ? launch_widths<T, true>(a, stream) : launch_widths<T, false>(a, stream);
}

template <typename T>
hipError_t resident_forms(int max_cols, size_t lds, int full, int *blocks) {
  return full ? resident_widths<T, true>(max_cols, lds, blocks)
              : resident_widths<T, false>(max_cols, lds, blocks);
}

}  // namespace

hipError_t launch_pipeline_nt128(int dtype, const JobLaunch &a, hipStream_t stream) {
  return dtype == 0 ? launch_forms<float>(a, stream) : launch_forms<double>(a, stream);
}

hipError_t resident_pipeline_nt128(int dtype, int max_cols, size_t lds, int full, int *blocks) {
  return dtype == 0 ? resident_forms<float>(max_cols, lds, full, blocks)
                    : resident_forms<double>(max_cols, lds, full, blocks);
}

}  // namespace pseg
