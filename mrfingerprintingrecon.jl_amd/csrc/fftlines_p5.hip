// Line-pass kernels for the lengths of part 5 (MSR_FFT_LINE_LENGTHS_P5).
#include "fftlines_impl.hpp"

namespace msr {

bool line_pass_part5(const LinePass& p, const void* tw, bool rows, hipStream_t st) {
    MSR_LINE_PART_BODY(MSR_FFT_LINE_LENGTHS_P5)
}

}  // namespace msr
