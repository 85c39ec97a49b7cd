// pipeline_kernel with 256 threads (4 wave64) per row: rows of up to
// 6144 samples, 4, 8, 12, 16, 24 samples per thread; each width in
// the demean / filter / AGC form, the form with every step but the IIR
// cascades, and the form with every step.
#include "pseg_pipeline.hpp"

namespace pseg {

namespace {

template <typename T, int FORM>
hipError_t launch_widths(const JobLaunch &a, hipStream_t stream) {
  const int per = (a.max_cols + 255) / 256;
  if (per <= 4) return launch_one<T, 256, 4, FORM>(a, stream);
  if (per <= 8) return launch_one<T, 256, 8, FORM>(a, stream);
  if (per <= 12) return launch_one<T, 256, 12, FORM>(a, stream);
  if (per <= 16) return launch_one<T, 256, 16, FORM>(a, stream);
  if (per <= 24) return launch_one<T, 256, 24, FORM>(a, stream);
  return hipErrorInvalidValue;
}

template <typename T, int FORM>
hipError_t resident_widths(int max_cols, size_t lds, int *blocks) {
  const int per = (max_cols + 255) / 256;
  if (per <= 4) return resident_one<T, 256, 4, FORM>(lds, blocks);
  if (per <= 8) return resident_one<T, 256, 8, FORM>(lds, blocks);
  if (per <= 12) return resident_one<T, 256, 12, FORM>(lds, blocks);
  if (per <= 16) return resident_one<T, 256, 16, FORM>(lds, blocks);
  if (per <= 24) return resident_one<T, 256, 24, FORM>(lds, blocks);
  return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_forms(const JobLaunch &a, hipStream_t stream) {
  return a.full == 2 ? launch_widths<T, 2>(a, stream)
         : a.full ? launch_widths<T, 1>(a, stream) : launch_widths<T, 0>(a, stream);
}

template <typename T>
hipError_t resident_forms(int max_cols, size_t lds, int full, int *blocks) {
  return full == 2 ? resident_widths<T, 2>(max_cols, lds, blocks)
         : full ? resident_widths<T, 1>(max_cols, lds, blocks)
                : resident_widths<T, 0>(max_cols, lds, blocks);
}

}  // namespace

hipError_t launch_pipeline_nt256(int dtype, const JobLaunch &a, hipStream_t stream) {
  return dtype == 0 ? launch_forms<float>(a, stream) : launch_forms<double>(a, stream);
}

hipError_t resident_pipeline_nt256(int dtype, int max_cols, size_t lds, int full, int *blocks) {
  return dtype == 0 ? resident_forms<float>(max_cols, lds, full, blocks)
                    : resident_forms<double>(max_cols, lds, full, blocks);
}

}  // namespace pseg
