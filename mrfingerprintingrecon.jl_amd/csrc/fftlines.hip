// Dispatch of the batched 1D FFT line passes (see fftlines.hpp) to the
// translation units that instantiate each length (fftlines_p*.hip).
#include "fftlines.hpp"
#include "fft.hpp"

namespace msr {

using LinePartFn = bool (*)(const LinePass&, const void*, bool, hipStream_t);
static const LinePartFn kLineParts[MSR_FFT_LINE_PARTS] = {
    line_pass_part0, line_pass_part1, line_pass_part2,
    line_pass_part3, line_pass_part4, line_pass_part5};

static void run_line_pass(const LinePass& p, msr_ctx* ctx, bool rows, hipStream_t st) {
    MSR_REQUIRE(fft_len_supported(p.nf), MSR_ERR_UNSUPPORTED, "no FFT kernel for length %d", p.nf);
    const void* tw = p.f64 ? (const void*)twiddle_table64(ctx, p.nf) : (const void*)twiddle_table(ctx, p.nf);
    for (LinePartFn f : kLineParts)
        if (f(p, tw, rows, st)) {
            MSR_CHECK_LAUNCH();
            return;
        }
    fail(MSR_ERR_UNSUPPORTED, "no FFT kernel for length %d", p.nf);
}

void run_rows(const LinePass& p, msr_ctx* ctx, hipStream_t st) { run_line_pass(p, ctx, true, st); }

void run_cols(const LinePass& p, msr_ctx* ctx, hipStream_t st) { run_line_pass(p, ctx, false, st); }

}  // namespace msr
