// Generic batched 1D FFT passes over lines of a grid, used by the
// backprojection and the Toeplitz-kernel construction:
//  - rows: lines contiguous in memory (x axis)
//  - cols: lines strided, adjacent lines adjacent in memory (y or z axis),
//          processed 8-16 lines per workgroup so loads and stores coalesce
// Each line has Lin valid inputs (zero-padded to the FFT length NF) and Lout
// selected outputs: SEL_FIRST bins 0..Lout-1, SEL_CENTER frequency
// r = j - Lout/2 (fftshift order), SEL_FFTORDER r = j or j - Lout (FFT order),
// each scaled by an optional real weight w[j] (deapodisation).  Inputs can be
// placed the same way (isel / win) for the forward (type-2) direction.
#pragma once

#include "internal.hpp"

namespace msr {

enum { SEL_FIRST = 0, SEL_CENTER = 1, SEL_FFTORDER = 2 };

struct LinePass {
    const void* in = nullptr;   // ComplexF32, or ComplexF64 when f64
    void* out = nullptr;
    int f64 = 0;                // arithmetic / data precision (the dense-kernel store rounds to FP32)
    int nf = 0;          // FFT length
    int dir = -1;        // -1 forward, +1 inverse
    int lin = 0, lout = 0, sel = SEL_FIRST;
    const float* w = nullptr;
    const double* w64 = nullptr;  // FP64 output weights (used instead of w when set)
    // input placement: input i goes to bin select(i, lin) scaled by win[i]
    // (SEL_FIRST: bins 0..lin-1, the rest zero; SEL_CENTER: frequency i - lin/2)
    int isel = SEL_FIRST;
    const float* win = nullptr;
    const double* win64 = nullptr;
    // rows: line l at in + l*in_pitch (contiguous), out + l*out_pitch
    int64_t nlines = 0, in_pitch = 0, out_pitch = 0;
    // cols: lines (b, col) with col < ncols adjacent in memory
    int64_t nb = 0, ncols = 0;
    int64_t in_bstride = 0, in_stride = 0, out_bstride = 0, out_stride = 0;
    // cols epilogue writing the dense Toeplitz kernel: b = packed pair p,
    // j = qy, col = qx:  lam[(qx*P + p)*ny2 + qy] = mask[qy*nx2 + qx] ? v : 0
    void* lam = nullptr;
    const uint8_t* mask = nullptr;
    int P = 0, lam_real = 0;
    int p0 = 0;                   // packed pair of batch line b is p0 + b
    int64_t nx2 = 0;
    // 3D variant of the dense store: col = qy*nx2 + qx, j = qz,
    // lam[((qy*nx2 + qx)*P + p)*nz2 + qz], mask[(qz*ny2 + qy)*nx2 + qx]
    int lam3d = 0;
    int64_t ny2 = 0;
    // device flag: the pass does nothing when *skip != 0 (early-stopped CG)
    const int* skip = nullptr;
    // cols passes with f64: input / output stored as cx64 (8 bytes per value,
    // common.hpp) instead of ComplexF64 (the 3D apply's compact intermediates)
    int in_compact = 0, out_compact = 0;
    // passes over 8-byte values (cx64, ComplexF32): one value per access
    // instead of two adjacent columns (or two consecutive line elements) per
    // 16-byte access (set by the launchers from MSR_COLS_SINGLE=1)
    int single = 0;
};

void run_rows(const LinePass& p, msr_ctx* ctx, hipStream_t st);
void run_cols(const LinePass& p, msr_ctx* ctx, hipStream_t st);

// per-length launchers, one per translation unit (fftlines_p*.hip): launch
// the rows (rows = true) or cols pass of p when p.nf is one of the part's
// lengths (tw: the length's twiddle table of p's precision) and return true
bool line_pass_part0(const LinePass& p, const void* tw, bool rows, hipStream_t st);
bool line_pass_part1(const LinePass& p, const void* tw, bool rows, hipStream_t st);
bool line_pass_part2(const LinePass& p, const void* tw, bool rows, hipStream_t st);
bool line_pass_part3(const LinePass& p, const void* tw, bool rows, hipStream_t st);
bool line_pass_part4(const LinePass& p, const void* tw, bool rows, hipStream_t st);
bool line_pass_part5(const LinePass& p, const void* tw, bool rows, hipStream_t st);

}  // namespace msr
