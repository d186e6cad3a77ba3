// Line-pass kernels for the lengths of part 4 (MSR_FFT_LINE_LENGTHS_P4).
#include "fftlines_impl.hpp"

namespace msr {

bool line_pass_part4(const LinePass& p, const void* tw, bool rows, hipStream_t st) {
    MSR_LINE_PART_BODY(MSR_FFT_LINE_LENGTHS_P4)
}

}  // namespace msr
