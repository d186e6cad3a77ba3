// Line-pass kernels for the lengths of part 0 (MSR_FFT_LINE_LENGTHS_P0).
#include "fftlines_impl.hpp"

namespace msr {

bool line_pass_part0(const LinePass& p, const void* tw, bool rows, hipStream_t st) {
    MSR_LINE_PART_BODY(MSR_FFT_LINE_LENGTHS_P0)
}

}  // namespace msr
