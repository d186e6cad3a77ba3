// Line-pass kernels for the lengths of part 1 (MSR_FFT_LINE_LENGTHS_P1).
#include "fftlines_impl.hpp"

namespace msr {

bool line_pass_part1(const LinePass& p, const void* tw, bool rows, hipStream_t st) {
    MSR_LINE_PART_BODY(MSR_FFT_LINE_LENGTHS_P1)
}

}  // namespace msr
